"""ctypes bindings for the CPU oracle (liboracle.so) and the reference harness
(_ref/libqmref.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The product never imports this.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libqmref.so")
TABLES = os.path.join(REPO, "open-qmin_amd", "stencils", "ldg_stencil_tables.json")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p
_i, _d, _ll = ctypes.c_int, ctypes.c_double, ctypes.c_int64

FORCE_FN = ctypes.CFUNCTYPE(None, _VP, _i, _D, _D)

# default phase constants of the openQmin CLI, rescaled as openQmin.cpp:173-175 does
PHASE_A, PHASE_B, PHASE_C = -0.172, -2.12, 1.73


def rescaled_phase(a=PHASE_A, b=PHASE_B, c=PHASE_C):
    return -1.0, -b / a, -c / a


def uniaxial_s0(A, B, C):
    """S0 = (-B + sqrt(B^2 - 24 A C)) / (6 C)  (openQmin.cpp:265)"""
    return (-B + np.sqrt(B * B - 24 * A * C)) / (6 * C)


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle not built: make -C oracle oracle")
        lib = ctypes.CDLL(ORACLE_LIB)
        _sig(lib, "orc_create", _VP, [_i, _i, _i])
        _sig(lib, "orc_destroy", None, [_VP])
        _sig(lib, "orc_set_threads", None, [_VP, _i])
        _sig(lib, "orc_neighbors", None, [_VP, _I])
        _sig(lib, "orc_get_types", None, [_VP, _I])
        _sig(lib, "orc_set_types", None, [_VP, _I])
        for n in ("orc_get_q", "orc_set_q", "orc_get_v", "orc_set_v", "orc_get_f"):
            _sig(lib, n, None, [_VP, _D])
        _sig(lib, "orc_random_init", None, [_VP, _i, _d])
        _sig(lib, "orc_random_uniform_init", None, [_VP, _i, _d])
        _sig(lib, "orc_director_init", None, [_VP, _D])
        _sig(lib, "orc_object_sites", _ll, [_VP, _i, _i, ctypes.POINTER(ctypes.c_int64)])
        _sig(lib, "orc_displace_object", _i, [_VP, _i, _i, _i])
        _sig(lib, "orc_object_force", _i, [_VP, _i, _D])
        _sig(lib, "orc_add_object", _i, [_VP, _I, _i, _D, _i, _d, _d])
        _sig(lib, "orc_wall", _i, [_VP, _i, _i, _i, _d, _d])
        _sig(lib, "orc_sphere", _i, [_VP, _d, _d, _d, _d, _i, _d, _d])
        _sig(lib, "orc_n_active", _ll, [_VP])
        _sig(lib, "orc_n_objects", _i, [_VP])
        _sig(lib, "orc_set_params", None, [_VP] + [_d] * 8 + [_i, _i])
        _sig(lib, "orc_set_efield", None, [_VP, _i] + [_d] * 6)
        _sig(lib, "orc_set_hfield", None, [_VP, _i] + [_d] * 6)
        _sig(lib, "orc_set_hfield_varying", None, [_VP, _i, _D] + [_d] * 3)
        _sig(lib, "orc_load_tables", None, [_i, _I, _I, _I, _D, _I, _I])
        _sig(lib, "orc_forces", None, [_VP, _D, _D])
        _sig(lib, "orc_first_derivatives", None, [_VP, _D, _D])
        _sig(lib, "orc_energy", _d, [_VP, _D, _D])
        _sig(lib, "orc_energy_density", None, [_VP, _D, _D])
        _sig(lib, "orc_stress_tensor", None, [_D, _d, _d, _D])
        _sig(lib, "orc_fire", _i, [_VP, _d, _d, _d, _d, _d, _d, _i, _d, _d, _i, _D])
        _sig(lib, "orc_adam", _i, [_VP, _d, _d, _d, _d, _d, _i, _D])
        _sig(lib, "orc_nesterov", _i, [_VP, _d, _d, _d, _i, _i, _D])
        _sig(lib, "orc_lolbfgs", _i, [_VP, _i, _d, _d, _d, _d, _i, _i, _D])
        _sig(lib, "orc_defects", None, [_VP, _i, _D])
        _sig(lib, "orc_eigenvalues", None, [_D, _D])
        load_tables(lib)
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_LIB)
        _sig(lib, "qmref_boundary_case", _i, [_I])
        _sig(lib, "qmref_bulk_l1", None, [_d, _D, _D, _D])
        _sig(lib, "qmref_boundary_l1", None, [_d, _D, _D, _I, _D])
        _sig(lib, "qmref_lk", None, [_i, _d, _i, _D, _D, _D, _D])
        _sig(lib, "qmref_anchoring_force", None, [_i, _d, _d, _D, _D, _D])
        _sig(lib, "qmref_anchoring_energy", _d, [_i, _d, _d, _D, _D])
        _sig(lib, "qmref_phase", None, [_D, _D, _D])
        _sig(lib, "qmref_eigenvalues", None, [_D, _D])
        _sig(lib, "qmref_determinant", _d, [_D])
        _sig(lib, "qmref_q_from_director", None, [_D, _d, _D])
        _sig(lib, "qmref_dotvec", _d, [_D, _D])
        _sig(lib, "qmref_dot", _d, [_D, _D])
        _sig(lib, "qmref_uniform_stream", None, [_i, _i, _D])
        _sig(lib, "qmref_model_create", _VP, [_i, _i, _i])
        _sig(lib, "qmref_model_destroy", None, [_VP])
        _sig(lib, "qmref_model_neighbors", None, [_VP, _I])
        _sig(lib, "qmref_model_random_init", None, [_VP, _i, _d])
        _sig(lib, "qmref_model_random_init_aligned", None, [_VP, _i, _d])
        _sig(lib, "qmref_model_object_sites", _i, [_VP, _i, _i, _I])
        _sig(lib, "qmref_model_displace_object", None, [_VP, _i, _i, _i])
        _sig(lib, "qmref_normalize_director", None, [_D])
        _sig(lib, "qmref_model_add_object", None, [_VP, _I, _i, _i, _d, _d])
        _sig(lib, "qmref_model_get_q", None, [_VP, _D])
        _sig(lib, "qmref_model_set_q", None, [_VP, _D])
        _sig(lib, "qmref_model_get_types", None, [_VP, _I])
        _sig(lib, "qmref_model_defects", None, [_VP, _i, _D])
        _sig(lib, "qmref_fire_run", _i, [_VP, _VP, _VP, _d, _d, _d, _d, _d, _d, _i, _d, _d, _i, _D])
        _sig(lib, "qmref_fire_minimize", _i, [_VP, _VP, _VP, _d, _d, _d, _d, _d, _d, _i, _d, _d, _i, _D])
        _sig(lib, "qmref_adam_run", _i, [_VP, _VP, _VP, _d, _d, _d, _d, _d, _i, _D])
        _sig(lib, "qmref_nesterov_run", _i, [_VP, _VP, _VP, _d, _d, _d, _i, _i, _D])
        _sig(lib, "qmref_lolbfgs_run", _i, [_VP, _VP, _VP, _i, _d, _d, _d, _d, _i, _i, _D])
        _ref = lib
    return _ref


def table_arrays(path=TABLES):
    """Flatten the stencil tables into (term, case slot, comp, coef, a, b) arrays."""
    with open(path) as fh:
        doc = json.load(fh)
    term, cslot, comp, coef, a, b = [], [], [], [], [], []
    for ti, name in enumerate(("L2", "L3", "L4", "L6")):
        for key, entries in doc["tables"][name].items():
            if key == "bulk":
                slot = 0
            else:
                code = int(key)
                slot = 1 + (code & 3) + 3 * ((code >> 2) & 3) + 9 * ((code >> 4) & 3)
            for e in entries:
                term.append(ti)
                cslot.append(slot)
                comp.append(e[0])
                coef.append(e[1])
                a.append(e[2])
                b.append(e[3] if len(e) == 4 else -1)
    return (np.array(term, np.int32), np.array(cslot, np.int32), np.array(comp, np.int32),
            np.array(coef, np.float64), np.array(a, np.int32), np.array(b, np.int32))


def load_tables(lib):
    arrs = table_arrays()
    lib.orc_load_tables(len(arrs[0]), iptr(arrs[0]), iptr(arrs[1]), iptr(arrs[2]), dptr(arrs[3]),
                        iptr(arrs[4]), iptr(arrs[5]))


class OracleSystem:
    """One lattice with its types, objects, Q/v/F and force parameters."""

    def __init__(self, lx, ly=None, lz=None):
        ly = lx if ly is None else ly
        lz = lx if lz is None else lz
        self.lib = oracle_lib()
        self.shape = (lx, ly, lz)
        self.n = lx * ly * lz
        self.h = self.lib.orc_create(lx, ly, lz)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def threads(self, n):
        self.lib.orc_set_threads(self.h, n)

    def neighbors(self):
        out = np.zeros(6 * self.n, np.int32)
        self.lib.orc_neighbors(self.h, iptr(out))
        return out.reshape(self.n, 6)

    def types(self):
        out = np.zeros(self.n, np.int32)
        self.lib.orc_get_types(self.h, iptr(out))
        return out

    def set_types(self, t):
        t = np.ascontiguousarray(t, np.int32)
        self.lib.orc_set_types(self.h, iptr(t))

    def q(self):
        out = np.zeros((self.n, 5))
        self.lib.orc_get_q(self.h, dptr(out))
        return out

    def v(self):
        out = np.zeros((self.n, 5))
        self.lib.orc_get_v(self.h, dptr(out))
        return out

    def f(self):
        out = np.zeros((self.n, 5))
        self.lib.orc_get_f(self.h, dptr(out))
        return out

    def set_q(self, q):
        q = np.ascontiguousarray(q, np.float64)
        self.lib.orc_set_q(self.h, dptr(q))

    def set_v(self, q):
        q = np.ascontiguousarray(q, np.float64)
        self.lib.orc_set_v(self.h, dptr(q))

    def random_init(self, seed, S0):
        self.lib.orc_random_init(self.h, seed, S0)

    def random_uniform_init(self, seed, S0):
        self.lib.orc_random_uniform_init(self.h, seed, S0)

    def director_init(self, n4):
        n4 = np.ascontiguousarray(n4, np.float64)
        assert n4.shape == (self.n, 4)
        self.lib.orc_director_init(self.h, dptr(n4))

    def object_sites(self, obj, surface=False):
        count = self.lib.orc_object_sites(self.h, obj, int(surface), None)
        assert count >= 0
        out = np.zeros(count, np.int64)
        self.lib.orc_object_sites(self.h, obj, int(surface), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out

    def displace_object(self, obj, direction, magnitude):
        assert self.lib.orc_displace_object(self.h, obj, direction, magnitude) == 0

    def object_force(self, obj):
        out = np.zeros(3)
        assert self.lib.orc_object_force(self.h, obj, dptr(out)) == 0
        return out

    def add_object(self, xyz, qb, btype, W, S0):
        xyz = np.ascontiguousarray(xyz, np.int32)
        qb = np.ascontiguousarray(qb, np.float64)
        return self.lib.orc_add_object(self.h, iptr(xyz), len(xyz), dptr(qb), btype, W, S0)

    def wall(self, xyz, plane, btype, W, S0):
        return self.lib.orc_wall(self.h, xyz, plane, btype, W, S0)

    def sphere(self, center, radius, btype, W, S0):
        return self.lib.orc_sphere(self.h, center[0], center[1], center[2], radius, btype, W, S0)

    def n_active(self):
        return self.lib.orc_n_active(self.h)

    def set_params(self, A, B, C, L1, L2=0.0, L3=0.0, L4=0.0, L6=0.0, multi=None, semantics=0):
        if multi is None:
            multi = int(L2 != 0 or L3 != 0 or L4 != 0 or L6 != 0)
        self.lib.orc_set_params(self.h, A, B, C, L1, L2, L3, L4, L6, int(multi), int(semantics))

    def set_efield(self, E, eps=1.0, eps0=1.0, deps=0.5, on=True):
        self.lib.orc_set_efield(self.h, int(on), E[0], E[1], E[2], eps, eps0, deps)

    def set_hfield(self, H, chi=1.0, mu0=1.0, dchi=0.5, on=True):
        self.lib.orc_set_hfield(self.h, int(on), H[0], H[1], H[2], chi, mu0, dchi)

    def set_hfield_varying(self, H3, chi=1.0, mu0=1.0, dchi=0.5, on=True):
        """per-site H field, array [n][3] in site order"""
        h = np.ascontiguousarray(H3, np.float64).reshape(self.n, 3)
        self._hvar = h
        self.lib.orc_set_hfield_varying(self.h, int(on), dptr(h), chi, mu0, dchi)

    def forces(self, q=None):
        q = self.q() if q is None else np.ascontiguousarray(q, np.float64)
        f = np.zeros((self.n, 5))
        self.lib.orc_forces(self.h, dptr(q), dptr(f))
        return f

    def first_derivatives(self, q=None):
        q = self.q() if q is None else np.ascontiguousarray(q, np.float64)
        d = np.zeros((self.n, 15))
        self.lib.orc_first_derivatives(self.h, dptr(q), dptr(d))
        return d

    def energy(self, q=None):
        q = self.q() if q is None else np.ascontiguousarray(q, np.float64)
        comps = np.zeros(5)
        total = self.lib.orc_energy(self.h, dptr(q), dptr(comps))
        return total, comps

    def energy_density(self, q=None):
        """Per-site energy density (landauDeGennesLC::energyDensity), site order."""
        q = self.q() if q is None else np.ascontiguousarray(q, np.float64)
        out = np.zeros(self.n)
        self.lib.orc_energy_density(self.h, dptr(q), dptr(out))
        return out

    def fire(self, dt=5e-4, alpha_start=0.99, dt_max=None, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9, n_min=4,
             fc=1e-12, alpha_min=0.0, max_iter=100):
        dt_max = 100 * dt if dt_max is None else dt_max
        trace = np.zeros((max_iter + 1, 6))
        it = self.lib.orc_fire(self.h, dt, alpha_start, dt_max, dt_inc, dt_dec, alpha_dec, n_min, fc, alpha_min,
                               max_iter, dptr(trace))
        return it, trace[:it]

    def adam(self, b1=0.9, b2=0.99, eps=1e-8, dt=1e-4, fc=1e-12, max_iter=100):
        trace = np.zeros(max_iter + 1)
        it = self.lib.orc_adam(self.h, b1, b2, eps, dt, fc, max_iter, dptr(trace))
        return it, trace[:it]

    def nesterov(self, dt=1e-4, mu=0.01, fc=1e-12, scheduled=False, max_iter=100):
        trace = np.zeros(max_iter + 1)
        it = self.lib.orc_nesterov(self.h, dt, mu, fc, int(scheduled), max_iter, dptr(trace))
        return it, trace[:it]

    def lolbfgs(self, m=5, dt=1e-4, c=1.0, fc=1e-12, tau=10.0, scheduled=False, max_iter=100):
        trace = np.zeros(max_iter + 1)
        it = self.lib.orc_lolbfgs(self.h, m, dt, c, fc, tau, int(scheduled), max_iter, dptr(trace))
        return it, trace[:it]

    def defects(self, kind=0):
        out = np.zeros(self.n)
        self.lib.orc_defects(self.h, kind, dptr(out))
        return out

    def force_callback(self):
        """(function pointer, context) for the reference harness minimizers."""
        return ctypes.cast(self.lib.orc_force_cb, _VP), ctypes.c_void_p(self.h)


class RefModel:
    """The reference's own qTensorLatticeModel (compiled from its sources)."""

    def __init__(self, lx, ly=None, lz=None):
        ly = lx if ly is None else ly
        lz = lx if lz is None else lz
        self.lib = ref_lib()
        self.n = lx * ly * lz
        self.h = self.lib.qmref_model_create(lx, ly, lz)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.qmref_model_destroy(self.h)
            self.h = None

    def neighbors(self):
        out = np.zeros(6 * self.n, np.int32)
        self.lib.qmref_model_neighbors(self.h, iptr(out))
        return out.reshape(self.n, 6)

    def random_init(self, seed, S0):
        self.lib.qmref_model_random_init(self.h, seed, S0)

    def random_init_aligned(self, seed, S0):
        self.lib.qmref_model_random_init_aligned(self.h, seed, S0)

    def object_sites(self, obj, surface=False):
        count = self.lib.qmref_model_object_sites(self.h, obj, int(surface), None)
        out = np.zeros(count, np.int32)
        self.lib.qmref_model_object_sites(self.h, obj, int(surface), iptr(out))
        return out.astype(np.int64)

    def displace_object(self, obj, direction, magnitude):
        self.lib.qmref_model_displace_object(self.h, obj, direction, magnitude)

    def add_object(self, sites, btype, P1, P2):
        sites = np.ascontiguousarray(sites, np.int32)
        self.lib.qmref_model_add_object(self.h, iptr(sites), len(sites), btype, P1, P2)

    def q(self):
        out = np.zeros((self.n, 5))
        self.lib.qmref_model_get_q(self.h, dptr(out))
        return out

    def set_q(self, q):
        q = np.ascontiguousarray(q, np.float64)
        self.lib.qmref_model_set_q(self.h, dptr(q))

    def types(self):
        out = np.zeros(self.n, np.int32)
        self.lib.qmref_model_get_types(self.h, iptr(out))
        return out

    def defects(self, kind=0):
        out = np.zeros(self.n)
        self.lib.qmref_model_defects(self.h, kind, dptr(out))
        return out

    def fire(self, force_fn, ctx, dt=5e-4, alpha_start=0.99, dt_max=None, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9,
             n_min=4, fc=1e-12, alpha_min=0.0, max_iter=100):
        dt_max = 100 * dt if dt_max is None else dt_max
        trace = np.zeros((max_iter + 1, 6))
        it = self.lib.qmref_fire_run(self.h, force_fn, ctx, dt, alpha_start, dt_max, dt_inc, dt_dec, alpha_dec,
                                     n_min, fc, alpha_min, max_iter, dptr(trace))
        return it, trace[:it]

    def adam(self, force_fn, ctx, b1=0.9, b2=0.99, eps=1e-8, dt=1e-4, fc=1e-12, max_iter=100):
        trace = np.zeros(max_iter + 1)
        it = self.lib.qmref_adam_run(self.h, force_fn, ctx, b1, b2, eps, dt, fc, max_iter, dptr(trace))
        return it, trace[:it]

    def nesterov(self, force_fn, ctx, dt=1e-4, mu=0.01, fc=1e-12, scheduled=False, max_iter=100):
        trace = np.zeros(max_iter + 1)
        it = self.lib.qmref_nesterov_run(self.h, force_fn, ctx, dt, mu, fc, int(scheduled), max_iter, dptr(trace))
        return it, trace[:it]

    def lolbfgs(self, force_fn, ctx, m=5, dt=1e-4, c=1.0, fc=1e-12, tau=10.0, scheduled=False, max_iter=100):
        trace = np.zeros(max_iter + 1)
        it = self.lib.qmref_lolbfgs_run(self.h, force_fn, ctx, m, dt, c, fc, tau, int(scheduled), max_iter,
                                        dptr(trace))
        return it, trace[:it]
