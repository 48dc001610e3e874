"""Python binding of libqmin.so (the HIP library) over its C ABI, include/qmin.h.

The classes mirror the reference's objects: a lattice model
(multirankQTensorLatticeModel on one rank), the landauDeGennesLC force computer
and the energyMinimizerFIRE / Adam / NesterovAG / LoLBFGS minimizers.  Every
call runs on the GPU; there is no CPU fallback.  A missing library raises.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libqmin.so")

QM_FP64, QM_FP32 = 0, 1
TILE_PER_LATTICE = -2   # QM_TILE_PER_LATTICE: an ensemble tile choice made per lattice
HOMEOTROPIC, DEGENERATE_PLANAR = 0, 1
SEMANTICS_REF_CPU, SEMANTICS_REF_GPU = 0, 1
FIELD_Q, FIELD_V, FIELD_F = 0, 1, 2

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_L = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p
_i, _d = ctypes.c_int, ctypes.c_double


class QminError(RuntimeError):
    pass


class FireState(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("nPositive", ctypes.c_int), ("converged", ctypes.c_int),
                ("nanDetected", ctypes.c_int), ("forceMax", ctypes.c_double), ("power", ctypes.c_double),
                ("alpha", ctypes.c_double), ("deltaT", ctypes.c_double), ("scaling", ctypes.c_double)]


# name -> argtypes (all return int status)
_API = {
    "qm_device_count": [ctypes.POINTER(ctypes.c_int)],
    "qm_lattice_create": [_i, _i, _i, _i, ctypes.POINTER(_VP)],
    "qm_lattice_destroy": [_VP],
    "qm_lattice_dims": [_VP] + [ctypes.POINTER(ctypes.c_int)] * 4,
    "qm_lattice_set_zchunk": [_VP, _i],
    "qm_lattice_set_tile": [_VP, _i],
    "qm_lattice_set_hfield": [_VP, _D],
    "qm_lattice_load_hfield": [_VP, ctypes.c_char_p],
    "qm_ldg_set_varying_hfield": [_VP, _i, _d, _d, _d],
    "qm_ws_conv_forward": [_VP, _i, _VP, _VP, _i, _VP, _VP, _VP, _i, _i, _i, _i, _i, _VP],
    "qm_ws_conv_wgrad": [_VP, _i, _VP, _i, _i, _i, _i, _i, _i, _VP, ctypes.c_longlong, _VP, _VP],
    "qm_ws_bias_grad": [_VP, ctypes.c_longlong, _i, _VP, _i, _VP, _VP],
    "qm_ws_adam": [_VP, _VP, _VP, _VP, ctypes.c_longlong, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                   ctypes.c_float, _i, _VP, _VP, _i, _i, _VP],
    "qm_ws_weights_to_bf16": [_VP, ctypes.c_longlong, _VP, _VP, _i, _i, _VP],
    "qm_ws_mse_grad": [_VP, _i, _VP, _VP, ctypes.c_longlong, ctypes.c_float, _D, _VP, _VP, _i, _VP],
    "qm_ws_features": [_VP, _d, _d, _d, _VP],
    "qm_ws_energy_grad": [_VP, _VP, ctypes.c_float, _VP, _i, _VP],
    "qm_ws_types": [_VP, _VP],
    "qm_ws_apply": [_VP, _VP, _i, _D],
    "qm_ws_extract_q": [_VP, _VP],
    "qm_ensemble_create": [ctypes.POINTER(_VP), _i, ctypes.POINTER(_VP)],
    "qm_ensemble_destroy": [_VP],
    "qm_ensemble_fire_minimize": [_VP, _VP, _VP, _i, ctypes.POINTER(ctypes.c_int), _D],
    "qm_ensemble_autotune": [_VP, _VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _D],
    "qm_ensemble_set_tile": [_VP, _i, _i],
    "qm_lattice_autotune": [_VP, _VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _D],
    "qm_lattice_set_precision": [_VP, _i],
    "qm_lattice_set_field": [_VP, _i, _D],
    "qm_lattice_get_field": [_VP, _i, _D],
    "qm_lattice_random_init": [_VP, _i, _d],
    "qm_lattice_uniform_init": [_VP, _d, _d, _d, _d],
    "qm_lattice_random_uniform_init": [_VP, _i, _d],
    "qm_lattice_director_init": [_VP, _D],
    "qm_lattice_add_object": [_VP, _I, ctypes.c_int64, _D, _i, _d, _d, ctypes.POINTER(ctypes.c_int)],
    "qm_lattice_wall": [_VP, _i, _i, _i, _d, _d, ctypes.POINTER(ctypes.c_int)],
    "qm_lattice_sphere": [_VP, _d, _d, _d, _d, _i, _d, _d, ctypes.POINTER(ctypes.c_int)],
    "qm_lattice_boundary_file": [_VP, ctypes.c_char_p],
    "qm_lattice_get_types": [_VP, _I],
    "qm_lattice_get_neighbors": [_VP, _I],
    "qm_lattice_n_active": [_VP, _L],
    "qm_lattice_n_boundary": [_VP, _L],
    "qm_lattice_defects": [_VP, _i, _D],
    "qm_lattice_save_state": [_VP, ctypes.c_char_p],
    "qm_lattice_object_sites": [_VP, _i, _i, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "qm_lattice_displace_object": [_VP, _i, _i, _i],
    "qm_ldg_object_force": [_VP, _VP, _i, _D],
    "qm_lattice_save_state_ex": [_VP, ctypes.c_char_p, _i, _i],
    "qm_lattice_load_state": [_VP, ctypes.c_char_p],
    "qm_lattice_copy_q": [_VP, _VP],
    "qm_ldg_create": [ctypes.POINTER(_VP)],
    "qm_ldg_destroy": [_VP],
    "qm_ldg_set_phase": [_VP, _d, _d, _d],
    "qm_ldg_set_elastic": [_VP, _d, _d, _d, _d, _d, _i],
    "qm_ldg_set_efield": [_VP, _i] + [_d] * 6,
    "qm_ldg_set_hfield": [_VP, _i] + [_d] * 6,
    "qm_ldg_set_semantics": [_VP, _i],
    "qm_ldg_compute_forces": [_VP, _VP],
    "qm_ldg_compute_energy": [_VP, _VP, _D, _D],
    "qm_ldg_energy_density": [_VP, _VP, _D],
    "qm_fire_create": [ctypes.POINTER(_VP)],
    "qm_fire_destroy": [_VP],
    "qm_fire_set_parameters": [_VP, _d, _d, _d, _d, _d, _d, _i, _d, _d],
    "qm_fire_set_max_iterations": [_VP, _i],
    "qm_fire_set_force_cutoff": [_VP, _d],
    "qm_fire_minimize": [_VP, _VP, _VP],
    "qm_fire_run_iterations": [_VP, _VP, _VP, _i],
    "qm_fire_run_iterations_noalloc": [_VP, _VP, _VP, _i],
    "qm_lattice_set_stream": [_VP, _VP],
    "qm_fire_get_state": [_VP, ctypes.POINTER(FireState)],
    "qm_fire_set_trace": [_VP, _i],
    "qm_fire_get_trace": [_VP, _D, _i, ctypes.POINTER(ctypes.c_int)],
    "qm_adam_create": [ctypes.POINTER(_VP)],
    "qm_adam_destroy": [_VP],
    "qm_adam_set_parameters": [_VP, _d, _d, _d, _d, _d],
    "qm_adam_set_max_iterations": [_VP, _i],
    "qm_adam_minimize": [_VP, _VP, _VP],
    "qm_adam_get_state": [_VP, ctypes.POINTER(ctypes.c_int), _D],
    "qm_adam_set_trace": [_VP, _i],
    "qm_adam_get_trace": [_VP, _D, _i, ctypes.POINTER(ctypes.c_int)],
    "qm_nesterov_set_trace": [_VP, _i],
    "qm_nesterov_get_trace": [_VP, _D, _i, ctypes.POINTER(ctypes.c_int)],
    "qm_lolbfgs_set_trace": [_VP, _i],
    "qm_lolbfgs_get_trace": [_VP, _D, _i, ctypes.POINTER(ctypes.c_int)],
    "qm_nesterov_create": [ctypes.POINTER(_VP)],
    "qm_nesterov_destroy": [_VP],
    "qm_nesterov_set_parameters": [_VP, _d, _d, _d, _i],
    "qm_nesterov_set_max_iterations": [_VP, _i],
    "qm_nesterov_minimize": [_VP, _VP, _VP],
    "qm_nesterov_get_state": [_VP, ctypes.POINTER(ctypes.c_int), _D],
    "qm_lolbfgs_create": [ctypes.POINTER(_VP)],
    "qm_lolbfgs_destroy": [_VP],
    "qm_lolbfgs_set_parameters": [_VP, _i, _d, _d, _d, _d, _i],
    "qm_lolbfgs_set_max_iterations": [_VP, _i],
    "qm_lolbfgs_minimize": [_VP, _VP, _VP],
    "qm_lolbfgs_get_state": [_VP, ctypes.POINTER(ctypes.c_int), _D],
    "qm_timing_enable": [_i],
    "qm_timing_reset": [],
    "qm_timing_get": [_i, _L, _D],
}

_lib = None
_torch_loaded_first = False   # torch's HIP runtime was in the process before the library


def library(path=LIB_PATH):
    """Load libqmin.so (raises if it has not been built)."""
    global _lib, _torch_loaded_first
    if _lib is None:
        _torch_loaded_first = "torch" in sys.modules
        if not os.path.exists(path):
            raise QminError("libqmin.so not built: run python open-qmin_amd/build.py")
        lib = ctypes.CDLL(path)
        lib.qm_last_error.restype = ctypes.c_char_p
        lib.qm_last_error.argtypes = []
        for name, args in _API.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib = lib
    return _lib


def check_torch_runtime_order(what):
    """torch must be imported before libqmin.so is loaded when the two share
    device memory or streams: torch ships a HIP runtime with the same soname as
    the system's, the first one loaded serves the process, and torch finds no
    GPU under the system runtime.  Raises QminError if the library came first."""
    if _lib is not None and not _torch_loaded_first:
        raise QminError("import torch before the first qmin call when %s "
                        "(one HIP runtime for torch and libqmin.so)" % what)


def exported_symbols():
    return ["qm_last_error"] + list(_API)


def _check(rc):
    if rc != 0:
        raise QminError("qmin error %d: %s" % (rc, _lib.qm_last_error().decode()))


def _dp(a):
    return a.ctypes.data_as(_D)


class Lattice:
    """One periodic lattice resident on the GPU (types, objects, Q, v, F)."""

    def __init__(self, lx, ly=None, lz=None, precision=QM_FP64):
        ly = lx if ly is None else ly
        lz = lx if lz is None else lz
        self.lib = library()
        self.shape = (lx, ly, lz)
        self.n = lx * ly * lz
        self.precision = precision
        h = _VP()
        _check(self.lib.qm_lattice_create(lx, ly, lz, precision, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.qm_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_zchunk(self, zchunk):
        _check(self.lib.qm_lattice_set_zchunk(self.h, zchunk))

    def set_stream(self, stream):
        """Run the lattice's work on a caller's HIP stream (an integer handle,
        e.g. torch.cuda.Stream().cuda_stream, default priority); None returns
        to the lattice's own stream.  A torch stream needs torch imported before
        the library was loaded (check_torch_runtime_order)."""
        if stream and "torch" in sys.modules:
            check_torch_runtime_order("a torch stream is passed to a lattice")
        _check(self.lib.qm_lattice_set_stream(self.h, ctypes.c_void_p(stream) if stream else None))

    def set_hfield(self, h3):
        """Spatially varying H field, [n][3] in site order (None removes it)."""
        if h3 is None:
            _check(self.lib.qm_lattice_set_hfield(self.h, None))
            return
        a = np.ascontiguousarray(h3, np.float64).reshape(self.n, 3)
        _check(self.lib.qm_lattice_set_hfield(self.h, _dp(a)))

    def load_hfield(self, path):
        _check(self.lib.qm_lattice_load_hfield(self.h, str(path).encode()))

    def set_tile(self, tile):
        """Tile shape of the fused pass: -1 automatic, 0 = 32x8, 1 = 64x4, 2 = 32x16, 3 = 64x8."""
        _check(self.lib.qm_lattice_set_tile(self.h, tile))

    def autotune(self, force):
        """Pick the fastest tile shape and z chunk of the fused FIRE pass for this
        lattice and force (timed on the current state, which is unchanged)."""
        t, z, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self.lib.qm_lattice_autotune(self.h, force.h, ctypes.byref(t), ctypes.byref(z), ctypes.byref(ms)))
        return t.value, z.value, ms.value

    def set_precision(self, precision):
        """Convert the device state to QM_FP64 or QM_FP32 (e.g. an FP64 polish)."""
        _check(self.lib.qm_lattice_set_precision(self.h, precision))
        self.precision = precision

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, np.float64).reshape(self.n, 5)
        _check(self.lib.qm_lattice_set_field(self.h, which, _dp(a)))

    def get_field(self, which):
        a = np.zeros((self.n, 5))
        _check(self.lib.qm_lattice_get_field(self.h, which, _dp(a)))
        return a

    def q(self):
        return self.get_field(FIELD_Q)

    def v(self):
        return self.get_field(FIELD_V)

    def f(self):
        return self.get_field(FIELD_F)

    def set_q(self, q):
        self.set_field(FIELD_Q, q)

    def random_init(self, seed, S0):
        _check(self.lib.qm_lattice_random_init(self.h, seed, S0))

    def uniform_init(self, director, S0):
        _check(self.lib.qm_lattice_uniform_init(self.h, director[0], director[1], director[2], S0))

    def random_uniform_init(self, seed, S0):
        """setNematicQTensorRandomly(noise, S0, globallyAligned=True): one director,
        from the first two draws of mt19937(seed), on every site with t <= 0."""
        _check(self.lib.qm_lattice_random_uniform_init(self.h, seed, S0))

    def director_init(self, directors):
        """setDirectorFromFunction: directors is an (n, 4) array of (nx, ny, nz, S0)
        in site order, or a function f(x, y, z) of integer coordinate arrays (each
        of n sites, x fastest) returning those four arrays."""
        if callable(directors):
            z, y, x = (c.ravel() for c in np.meshgrid(np.arange(self.shape[2]), np.arange(self.shape[1]),
                                                      np.arange(self.shape[0]), indexing="ij"))
            directors = np.stack([np.broadcast_to(np.asarray(a, np.float64), x.shape) for a in directors(x, y, z)],
                                 axis=1)
        d = np.ascontiguousarray(directors, np.float64)
        if d.shape != (self.n, 4):
            raise ValueError("directors must have shape (n, 4)")
        _check(self.lib.qm_lattice_director_init(self.h, _dp(d)))

    def add_object(self, xyz, qb, btype, W, S0):
        xyz = np.ascontiguousarray(xyz, np.int32)
        qb = np.ascontiguousarray(qb, np.float64)
        j = ctypes.c_int()
        _check(self.lib.qm_lattice_add_object(self.h, xyz.ctypes.data_as(_I), len(xyz), _dp(qb), btype, W, S0,
                                              ctypes.byref(j)))
        return j.value

    def wall(self, xyz, plane, btype, W, S0):
        j = ctypes.c_int()
        _check(self.lib.qm_lattice_wall(self.h, xyz, plane, btype, W, S0, ctypes.byref(j)))
        return j.value

    def sphere(self, center, radius, btype, W, S0):
        j = ctypes.c_int()
        _check(self.lib.qm_lattice_sphere(self.h, center[0], center[1], center[2], radius, btype, W, S0,
                                          ctypes.byref(j)))
        return j.value

    def boundary_file(self, path):
        _check(self.lib.qm_lattice_boundary_file(self.h, path.encode()))

    def types(self):
        t = np.zeros(self.n, np.int32)
        _check(self.lib.qm_lattice_get_types(self.h, t.ctypes.data_as(_I)))
        return t

    def neighbors(self):
        nb = np.zeros((self.n, 6), np.int32)
        _check(self.lib.qm_lattice_get_neighbors(self.h, nb.ctypes.data_as(_I)))
        return nb

    def n_active(self):
        v = ctypes.c_int64()
        _check(self.lib.qm_lattice_n_active(self.h, ctypes.byref(v)))
        return v.value

    def n_boundary(self):
        v = ctypes.c_int64()
        _check(self.lib.qm_lattice_n_boundary(self.h, ctypes.byref(v)))
        return v.value

    def defects(self, kind=0):
        out = np.zeros(self.n)
        _check(self.lib.qm_lattice_defects(self.h, kind, _dp(out)))
        return out

    def object_sites(self, obj, surface=False):
        """Sites of object obj (0-based): its own in creation order, or its
        surface sites in order of discovery (cubicLattice::boundarySites /
        surfaceSites)."""
        count = ctypes.c_int64()
        _check(self.lib.qm_lattice_object_sites(self.h, obj, int(surface), None, ctypes.byref(count)))
        out = np.zeros(count.value, np.int64)
        _check(self.lib.qm_lattice_object_sites(self.h, obj, int(surface),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                ctypes.byref(count)))
        return out

    def displace_object(self, obj, direction, magnitude=1):
        """cubicLattice::displaceBoundaryObject: direction 0..5 = x-, x+, y-, y+, z-, z+."""
        _check(self.lib.qm_lattice_displace_object(self.h, obj, direction, magnitude))

    def save_state(self, path, stride=1, defect_type=0):
        """saveState's text format (one line per site whose coordinates are all
        multiples of stride; last column computeDefectMeasures(defect_type))."""
        _check(self.lib.qm_lattice_save_state_ex(self.h, path.encode(), stride, defect_type))

    def load_state(self, path):
        _check(self.lib.qm_lattice_load_state(self.h, path.encode()))


class LandauDeGennesLC:
    """landauDeGennesLC: phase, L1..L6 distortion, anchoring and field terms."""

    def __init__(self, A=-1.0, B=0.0, C=0.0, L1=2.32, L2=0.0, L3=0.0, L4=0.0, L6=0.0, multi=None,
                 semantics=SEMANTICS_REF_GPU):
        self.lib = library()
        h = _VP()
        _check(self.lib.qm_ldg_create(ctypes.byref(h)))
        self.h = h
        self.set_phase(A, B, C)
        self.set_elastic(L1, L2, L3, L4, L6, multi)
        self.set_semantics(semantics)

    def close(self):
        if getattr(self, "h", None):
            self.lib.qm_ldg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_phase(self, A, B, C):
        _check(self.lib.qm_ldg_set_phase(self.h, A, B, C))

    def set_elastic(self, L1, L2=0.0, L3=0.0, L4=0.0, L6=0.0, multi=None):
        if multi is None:
            multi = int(L2 != 0 or L3 != 0 or L4 != 0 or L6 != 0)
        _check(self.lib.qm_ldg_set_elastic(self.h, L1, L2, L3, L4, L6, int(multi)))

    def set_efield(self, E, eps=1.0, eps0=1.0, deps=0.5, on=True):
        _check(self.lib.qm_ldg_set_efield(self.h, int(on), E[0], E[1], E[2], eps, eps0, deps))

    def set_hfield(self, H, chi=1.0, mu0=1.0, dchi=0.5, on=True):
        _check(self.lib.qm_ldg_set_hfield(self.h, int(on), H[0], H[1], H[2], chi, mu0, dchi))

    def set_varying_hfield(self, chi=1.0, mu0=1.0, dchi=0.5, on=True):
        """Add the lattice's spatially varying H field (Lattice.set_hfield / load_hfield)."""
        _check(self.lib.qm_ldg_set_varying_hfield(self.h, int(on), chi, mu0, dchi))

    def set_semantics(self, s):
        _check(self.lib.qm_ldg_set_semantics(self.h, s))

    def compute_forces(self, lattice):
        _check(self.lib.qm_ldg_compute_forces(self.h, lattice.h))
        return lattice.f()

    def compute_energy(self, lattice):
        tot = ctypes.c_double()
        comps = np.zeros(5)
        _check(self.lib.qm_ldg_compute_energy(self.h, lattice.h, ctypes.byref(tot), _dp(comps)))
        return tot.value, comps

    def energy_density(self, lattice):
        """landauDeGennesLC::energyDensity: per-site energy, site order (0 on
        object sites); compute_energy's total is its sum."""
        out = np.zeros(lattice.n)
        _check(self.lib.qm_ldg_energy_density(self.h, lattice.h, _dp(out)))
        return out

    def object_force(self, lattice, obj):
        """computeObjectForces: the stress-tensor force on object obj (0-based);
        one-constant models only, as in the reference."""
        out = np.zeros(3)
        _check(self.lib.qm_ldg_object_force(self.h, lattice.h, obj, _dp(out)))
        return out


class FIRE:
    """energyMinimizerFIRE on the device."""

    def __init__(self, dt=5e-4, alpha_start=0.99, dt_max=None, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9, n_min=4,
                 fc=1e-12, alpha_min=0.0, max_iter=100):
        self.lib = library()
        h = _VP()
        _check(self.lib.qm_fire_create(ctypes.byref(h)))
        self.h = h
        dt_max = 100 * dt if dt_max is None else dt_max
        _check(self.lib.qm_fire_set_parameters(self.h, dt, alpha_start, dt_max, dt_inc, dt_dec, alpha_dec, n_min, fc,
                                               alpha_min))
        self.set_max_iterations(max_iter)

    def close(self):
        if getattr(self, "h", None):
            self.lib.qm_fire_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_max_iterations(self, n):
        _check(self.lib.qm_fire_set_max_iterations(self.h, n))

    def set_force_cutoff(self, fc):
        """energyMinimizerFIRE::setForceCutoff; dt, alpha and N+ are kept."""
        _check(self.lib.qm_fire_set_force_cutoff(self.h, fc))

    def set_trace(self, capacity):
        _check(self.lib.qm_fire_set_trace(self.h, capacity))

    def trace(self, cap=1 << 20):
        st = self.state()
        cap = min(cap, max(st.iterations, 1))
        rows = np.zeros((cap, 6))
        n = ctypes.c_int()
        _check(self.lib.qm_fire_get_trace(self.h, _dp(rows), cap, ctypes.byref(n)))
        return rows[:n.value]

    def minimize(self, force, lattice):
        _check(self.lib.qm_fire_minimize(self.h, force.h, lattice.h))
        return self.state()

    def run_iterations(self, force, lattice, n):
        _check(self.lib.qm_fire_run_iterations(self.h, force.h, lattice.h, n))
        return self.state()

    def run_iterations_noalloc(self, force, lattice, n):
        """run_iterations with no device allocation (partial sums in the
        library's fixed scratch array; no trace)."""
        _check(self.lib.qm_fire_run_iterations_noalloc(self.h, force.h, lattice.h, n))
        return self.state()

    def state(self):
        s = FireState()
        _check(self.lib.qm_fire_get_state(self.h, ctypes.byref(s)))
        return s


class _SimpleMinimizer:
    _prefix = None

    def __init__(self):
        self.lib = library()
        h = _VP()
        _check(getattr(self.lib, self._prefix + "_create")(ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, self._prefix + "_destroy")(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_max_iterations(self, n):
        _check(getattr(self.lib, self._prefix + "_set_max_iterations")(self.h, n))

    def minimize(self, force, lattice):
        _check(getattr(self.lib, self._prefix + "_minimize")(self.h, force.h, lattice.h))
        return self.state()

    def state(self):
        it = ctypes.c_int()
        fm = ctypes.c_double()
        _check(getattr(self.lib, self._prefix + "_get_state")(self.h, ctypes.byref(it), ctypes.byref(fm)))
        return it.value, fm.value

    def set_trace(self, capacity):
        """Record forceMax of every iteration (up to capacity)."""
        self._trace_cap = capacity
        _check(getattr(self.lib, self._prefix + "_set_trace")(self.h, capacity))

    def trace(self):
        cap = getattr(self, "_trace_cap", 0)
        out = np.zeros(max(cap, 1))
        n = ctypes.c_int()
        _check(getattr(self.lib, self._prefix + "_get_trace")(
            self.h, out.ctypes.data_as(_D), cap, ctypes.byref(n)))
        return out[:n.value].copy()


class Adam(_SimpleMinimizer):
    """energyMinimizerAdam on the device."""
    _prefix = "qm_adam"

    def __init__(self, beta1=0.9, beta2=0.99, eps=1e-8, dt=1e-4, fc=1e-12, max_iter=1000):
        super().__init__()
        _check(self.lib.qm_adam_set_parameters(self.h, beta1, beta2, eps, dt, fc))
        self.set_max_iterations(max_iter)


class Nesterov(_SimpleMinimizer):
    """energyMinimizerNesterovAG on the device."""
    _prefix = "qm_nesterov"

    def __init__(self, dt=1e-4, mu=0.01, fc=1e-12, scheduled=False, max_iter=1000):
        super().__init__()
        _check(self.lib.qm_nesterov_set_parameters(self.h, dt, mu, fc, int(scheduled)))
        self.set_max_iterations(max_iter)


class LoLBFGS(_SimpleMinimizer):
    """energyMinimizerLoLBFGS on the device."""
    _prefix = "qm_lolbfgs"

    def __init__(self, m=5, dt=1e-4, c=1.0, fc=1e-12, tau=10.0, scheduled=False, max_iter=1000):
        super().__init__()
        _check(self.lib.qm_lolbfgs_set_parameters(self.h, m, dt, c, fc, tau, int(scheduled)))
        self.set_max_iterations(max_iter)


class Ensemble:
    """Many lattices (one precision) minimized together: one launch per kernel
    over all of them, hipGraph-replayed batches of iterations."""

    def __init__(self, lattices):
        self.lib = library()
        self.lattices = list(lattices)      # keep them alive as long as the ensemble
        arr = (_VP * len(self.lattices))(*[l.h for l in self.lattices])
        h = _VP()
        _check(self.lib.qm_ensemble_create(arr, len(self.lattices), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.qm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def fire_minimize(self, fire, force, graph_iterations=16):
        """FIRE with fire's parameters on every lattice; returns per-lattice
        (iterations, forceMax) arrays."""
        n = len(self.lattices)
        it = (ctypes.c_int * n)()
        fm = np.zeros(n)
        _check(self.lib.qm_ensemble_fire_minimize(self.h, fire.h, force.h, graph_iterations, it, _dp(fm)))
        return np.array(list(it)), fm

    def autotune(self, force):
        """Time the batched pass for every tile shape and z chunk on the current
        states (left unchanged) and keep the fastest; returns (tile, zchunk, ms)."""
        tile, zc, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self.lib.qm_ensemble_autotune(self.h, force.h, ctypes.byref(tile), ctypes.byref(zc),
                                             ctypes.byref(ms)))
        return tile.value, zc.value, ms.value

    def set_tile(self, tile, zchunk=0):
        """The batched pass's tile (-1 default, TILE_PER_LATTICE, or 0..3) and z
        chunk (0: automatic) for later fire_minimize calls."""
        _check(self.lib.qm_ensemble_set_tile(self.h, tile, zchunk))


def timing_enable(on=True):
    _check(library().qm_timing_enable(int(on)))


def timing_reset():
    _check(library().qm_timing_reset())


def timing_get(kernel):
    n = ctypes.c_int64()
    ms = ctypes.c_double()
    _check(library().qm_timing_get(kernel, ctypes.byref(n), ctypes.byref(ms)))
    return n.value, ms.value


def device_count():
    c = ctypes.c_int()
    _check(library().qm_device_count(ctypes.byref(c)))
    return c.value
