"""The CPU oracle against the reference's own outputs.

Fixtures in tests/golden/reference_vectors.npz come from the reference code
compiled from its sources (tests/golden/make_golden.py); where the compiled
reference (oracle/_ref) is present the same checks also run live.
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import pyoracle as po
from conftest import REPO

CASES = [xx + 4 * yy + 16 * zz for zz in range(3) for yy in range(3) for xx in range(3)]
live_ref = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built (needs the reference checkout)")


def _eval_table(entries, L, x):
    out = np.zeros(5)
    for e in entries:
        if len(e) == 3:
            out[e[0]] += e[1] * x[e[2]]
        else:
            out[e[0]] += e[1] * x[e[2]] * x[e[3]]
    return L * out


@pytest.fixture(scope="module")
def tables():
    with open(po.TABLES) as fh:
        return json.load(fh)["tables"]


def test_stencil_tables_reproduce_reference_outputs(golden, tables):
    X = golden["stencil_x"]
    for k in (2, 3, 4, 6):
        for c in ["bulk"] + [str(c) for c in CASES]:
            ref = golden["stencil_L%d_%s" % (k, c)]
            for i in range(len(X)):
                got = _eval_table(tables["L%d" % k][c], 1.7, X[i])
                np.testing.assert_allclose(got, ref[i], rtol=1e-13, atol=1e-12 * max(1, np.abs(ref[i]).max()),
                                           err_msg="L%d case %s" % (k, c))


@live_ref
def test_stencil_tables_match_live_reference(tables):
    ref = po.ref_lib()
    rng = np.random.default_rng(99)
    for k in (2, 3, 4, 6):
        for c in [-1] + CASES:
            x = rng.standard_normal(125)
            out = np.zeros(5)
            ref.qmref_lk(k, 0.9, c, po.dptr(x[:5]), po.dptr(x[5:35]), po.dptr(x[35:]), po.dptr(out))
            got = _eval_table(tables["L%d" % k]["bulk" if c < 0 else str(c)], 0.9, x)
            np.testing.assert_allclose(got, out, rtol=1e-13, atol=1e-12 * max(1, np.abs(out).max()))


def test_boundary_case_code_covers_27_cases(tables):
    for k in (2, 3, 4, 6):
        assert sorted(int(c) for c in tables["L%d" % k] if c != "bulk") == sorted(CASES)
        # the bulk form is the no-object case (lcForces.h:90-97 vs :66-80)
        assert tables["L%d" % k]["bulk"] == tables["L%d" % k]["0"]


def test_phase_parts(golden):
    lib = po.oracle_lib()
    lib.orc_phase_parts.argtypes = [po._D, po._D, po._D]
    for q, g, t in zip(golden["phase_q"], golden["phase_grad"], golden["phase_tr"]):
        q = np.ascontiguousarray(q)
        og, ot = np.zeros(15), np.zeros(3)
        lib.orc_phase_parts(po.dptr(q), po.dptr(og), po.dptr(ot))
        np.testing.assert_allclose(og, g, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(ot, t, rtol=1e-13, atol=1e-15)


def test_l1_stencils(golden):
    lib = po.oracle_lib()
    lib.orc_l1_site.argtypes = [ctypes.c_double, ctypes.c_int, po._D, po._D, po._I, po._D]
    X, T = golden["stencil_x"], golden["l1_types"]
    for i in range(len(X)):
        x, t = np.ascontiguousarray(X[i]), np.ascontiguousarray(T[i])
        out = np.zeros(5)
        lib.orc_l1_site(2.32, 0, po.dptr(x[:5]), po.dptr(x[5:35]), po.iptr(t), po.dptr(out))
        np.testing.assert_allclose(out, golden["l1_bulk"][i], rtol=1e-13, atol=1e-13)
        lib.orc_l1_site(2.32, 1, po.dptr(x[:5]), po.dptr(x[5:35]), po.iptr(t), po.dptr(out))
        np.testing.assert_allclose(out, golden["l1_boundary"][i], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kind,bt", [("h", 0), ("p", 1)])
def test_anchoring(golden, kind, bt):
    lib = po.oracle_lib()
    lib.orc_anchoring_force.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, po._D, po._D, po._D]
    lib.orc_anchoring_energy.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, po._D, po._D]
    lib.orc_anchoring_energy.restype = ctypes.c_double
    for i, q in enumerate(golden["anch_q"]):
        q = np.ascontiguousarray(q)
        qb = np.ascontiguousarray(golden["anch_qb_" + kind][i])
        f = np.zeros(5)
        lib.orc_anchoring_force(bt, 5.3, 0.53, po.dptr(q), po.dptr(qb), po.dptr(f))
        ref = golden["anch_force_" + kind][i]
        np.testing.assert_allclose(f, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        e = lib.orc_anchoring_energy(bt, 5.3, 0.53, po.dptr(q), po.dptr(qb))
        np.testing.assert_allclose(e, golden["anch_energy_" + kind][i], rtol=1e-12)


def test_eigenvalues(golden):
    lib = po.oracle_lib()
    for q, ev in zip(golden["eig_q"], golden["eig_ev"]):
        out = np.zeros(3)
        lib.orc_eigenvalues(po.dptr(np.ascontiguousarray(q)), po.dptr(out))
        np.testing.assert_allclose(out, ev, rtol=1e-12, atol=1e-14)


def test_lattice_init_types_defects(golden, phase):
    A, B, C, S0 = phase
    shape = tuple(golden["small_shape"])
    o = po.OracleSystem(*shape)
    assert np.array_equal(o.neighbors(), golden["small_neighbors"])
    o.random_init(1234, S0)
    assert np.array_equal(o.q(), golden["small_init_q"]), "seeded initial Q must be bit-exact"
    o.wall(2, 0, 1, 5.3, S0)
    o.sphere((3.0, 2.5, 3.5), 1.8, 0, 5.3, S0)
    assert np.array_equal(o.types(), golden["small_types"])
    d = o.defects(0)
    lc = o.types() <= 0
    np.testing.assert_allclose(d[lc], golden["small_defects"][lc], rtol=1e-13, atol=1e-15)


@live_ref
@pytest.mark.parametrize("seed", [1234, 7])
def test_random_uniform_init_matches_reference(phase, seed):
    """openQmin -z 1: setNematicQTensorRandomly(noise, S0, globallyAligned=true)
    of the compiled reference, bit for bit."""
    A, B, C, S0 = phase
    r = po.RefModel(6, 5, 4)
    r.random_init_aligned(seed, S0)
    o = po.OracleSystem(6, 5, 4)
    o.random_uniform_init(seed, S0)
    assert np.array_equal(o.q(), r.q())


def test_random_uniform_director_is_the_streams_first_two_draws(phase):
    """The aligned texture is uniaxial with order S0 along the director of the
    first two doubles of the seeded stream (theta = acos(2u - 1), phi = 2 pi u):
    checked through the eigen-decomposition of Q, not the Q formula."""
    import math
    A, B, C, S0 = phase
    raw = np.random.RandomState(99)._bit_generator.random_raw(4).astype(np.float64)
    u = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64
    theta, phi = math.acos(2 * u[0] - 1), 2 * math.pi * u[1]
    n = np.array([math.cos(phi) * math.sin(theta), math.sin(phi) * math.sin(theta), math.cos(theta)])
    o = po.OracleSystem(4, 3, 2)
    o.random_uniform_init(99, S0)
    q = o.q()
    assert np.array_equal(q, np.broadcast_to(q[0], q.shape))
    Q = np.array([[q[0, 0], q[0, 1], q[0, 2]], [q[0, 1], q[0, 3], q[0, 4]], [q[0, 2], q[0, 4], -q[0, 0] - q[0, 3]]])
    w, v = np.linalg.eigh(Q)
    assert abs(abs(v[:, 2] @ n) - 1) < 1e-14
    np.testing.assert_allclose(w, [-S0 / 2, -S0 / 2, S0], atol=1e-14)


@live_ref
def test_director_function_init_matches_reference(phase):
    """openQmin -z 3 (setDirectorFromFunction with the example function of
    setInitialConditions.h): the test helper's normalisation equals the
    reference's normalizeDirector, and the oracle's Q equals the reference's
    qTensorFromDirector at every site, bit for bit."""
    from conftest import example_directors
    shape = (6, 5, 4)
    n4 = example_directors(shape)
    lib = po.ref_lib()
    i = 0
    for z in range(shape[2]):
        for y in range(shape[1]):
            for x in range(shape[0]):
                import math
                v = np.array([float(x), x + y + math.cos(z), 0.0])
                lib.qmref_normalize_director(po.dptr(v))
                assert np.array_equal(v, n4[i, :3])
                i += 1
    o = po.OracleSystem(*shape)
    o.director_init(n4)
    ref = np.zeros((len(n4), 5))
    for i in range(len(n4)):
        d = np.ascontiguousarray(n4[i, :3])
        lib.qmref_q_from_director(po.dptr(d), n4[i, 3], po.dptr(ref[i]))
    assert np.array_equal(o.q(), ref)


def test_uniform_stream_is_libstdcxx_generate_canonical(golden):
    """getRealUniform = uniform_real_distribution<double> over mt19937: two 32-bit
    draws per double, (a + b 2^32) / 2^64 (libstdc++ generate_canonical).  numpy's
    legacy RandomState(seed) seeds MT19937 with the same init_genrand."""
    bg = np.random.RandomState(1234)._bit_generator
    raw = bg.random_raw(128).astype(np.float64)
    u = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64
    np.testing.assert_array_equal(u, golden["uniform_1234"])


def test_mt19937_known_answer():
    # The C++ standard fixes the 10000th output of a default-seeded mt19937.
    bg = np.random.RandomState(5489)._bit_generator
    assert int(bg.random_raw(10000)[-1]) == 4123659995


def _scene(phase, params, sem, ef=None):
    A, B, C, S0 = phase
    o = po.OracleSystem(6, 5, 7)
    o.random_init(1234, S0)
    o.wall(2, 0, 1, 5.3, S0)
    o.sphere((3.0, 2.5, 3.5), 1.8, 0, 5.3, S0)
    o.set_params(A, B, C, *params, semantics=sem)
    if ef is not None:
        o.set_efield(ef, 1.0, 1.0, 1.0)
    return o


@pytest.mark.parametrize("name,params,sem,ef", [
    ("one", (2.32, 0.0, 0.0, 0.0, 0.0), 0, None),
    ("multi", (2.32, 0.5, 0.3, 0.0, 0.0), 0, None),
    ("all", (2.32, 1.0, 0.5, 0.2, 0.3), 0, (0.0, 0.0, 0.5)),
    ("multi_gpu", (2.32, 0.5, 0.3, 0.0, 0.0), 1, None),
])
def test_fire_trajectory_bit_exact(golden, phase, name, params, sem, ef):
    o = _scene(phase, params, sem, ef)
    it, tr = o.fire(max_iter=40)
    assert it == 40
    assert np.array_equal(tr, golden["fire_%s_trace" % name])
    assert np.array_equal(o.q(), golden["fire_%s_q" % name])


def test_adam_nesterov_lolbfgs_bit_exact(golden, phase):
    p = (2.32, 0.5, 0.3, 0.0, 0.0)
    o = _scene(phase, p, 0)
    it, tr = o.adam(dt=1e-3, max_iter=30)
    assert np.array_equal(tr, golden["adam_trace"]) and np.array_equal(o.q(), golden["adam_q"])
    o = _scene(phase, p, 0)
    it, tr = o.nesterov(dt=1e-2, mu=0.5, scheduled=True, max_iter=30)
    assert np.array_equal(tr, golden["nesterov_trace"]) and np.array_equal(o.q(), golden["nesterov_q"])
    o = _scene(phase, p, 0)
    it, tr = o.lolbfgs(m=8, dt=1e-2, max_iter=30)
    assert np.array_equal(tr, golden["lolbfgs_trace"]) and np.array_equal(o.q(), golden["lolbfgs_q"])


@live_ref
def test_fire_minimize_loop_matches_reference_minimize(phase):
    """The reference's own minimize() (stopping rule included) against the oracle."""
    A, B, C, S0 = phase
    o = _scene(phase, (2.32, 0.0, 0.0, 0.0, 0.0), 0)
    r = po.RefModel(6, 5, 7)
    r.set_q(o.q())
    from golden.make_golden import sphere_sites  # noqa: E402
    wall = np.array([x + 6 * (y + 5 * 0) for x in range(6) for y in range(5)], np.int32)
    r.add_object(wall, 1, 5.3, S0)
    r.add_object(sphere_sites((6, 5, 7), (3.0, 2.5, 3.5), 1.8), 0, 5.3, S0)
    fn, ctx = o.force_callback()
    fmax = np.zeros(1)
    lib = po.ref_lib()
    its = lib.qmref_fire_minimize(r.h, fn, ctx, 5e-4, 0.99, 0.05, 1.1, 0.95, 0.9, 4, 0.05, 0.0, 5000, po.dptr(fmax))
    o2 = _scene(phase, (2.32, 0.0, 0.0, 0.0, 0.0), 0)
    it2, tr = o2.fire(fc=0.05, max_iter=5000)
    assert its == it2 and tr[-1, 0] == fmax[0] and tr[-1, 0] <= 0.05
    assert np.array_equal(r.q(), o2.q())


REF_ENERGY_SRC = "/root/reference/src/forces/landauDeGennesLC.cpp"


@pytest.mark.skipif(not os.path.exists(REF_ENERGY_SRC), reason="reference checkout absent")
def test_distortion_energy_matches_reference_polynomials():
    """The oracle writes the L1..L6 energy densities from tensor invariants; the
    reference spells them out as polynomials (landauDeGennesLC.cpp:268-335).
    Evaluate the reference's own expressions, read from its source text, on
    random inputs and compare term by term."""
    import re
    src = open(REF_ENERGY_SRC).read()
    body = src[src.index("void landauDeGennesLC::computeEnergyCPU"):src.index("#ifdef ENABLE_CUDA\n\nvoid landauDeGennesLC::computeForceGPU")]
    exprs = {}
    for name in ("L1", "L2", "L3", "L4", "L6"):
        m = re.search(r"if\(%s != 0 \)\s*\{\s*distortionEnergyAtSite \+= (.*?);\s*\};" % name, body, re.S)
        assert m, name
        e = m.group(1)
        e = re.sub(r"firstDerivative([XYZ])\[(\d)\]", r"d\1[\2]", e).replace("qCurrent", "q")
        exprs[name] = compile(e, name, "eval")
    lib = po.oracle_lib()
    lib.orc_distortion_density.argtypes = [po._D] * 5
    rng = np.random.default_rng(3)
    for _ in range(50):
        q, dX, dY, dZ = (np.ascontiguousarray(rng.standard_normal(5)) for _ in range(4))
        e = np.zeros(5)
        lib.orc_distortion_density(po.dptr(q), po.dptr(dX), po.dptr(dY), po.dptr(dZ), po.dptr(e))
        env = {"q": q, "dX": dX, "dY": dY, "dZ": dZ}
        for i, name in enumerate(("L1", "L2", "L3", "L4", "L6")):
            ref = eval(exprs[name], {name: 1.0}, env)
            np.testing.assert_allclose(e[i], ref, rtol=1e-12, atol=1e-12, err_msg=name)


def test_varying_field_oracle_reduces_to_uniform_and_is_local():
    """The oracle's spatially varying H field: a constant field equals the
    uniform H field (whose force and energy are pinned against the reference's
    source expressions and by finite differences in test_oracle_assembly.py),
    and a change at one site changes only that site's force."""
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    shape = (9, 8, 7)
    n = shape[0] * shape[1] * shape[2]
    H = (0.3, -0.4, 0.7)
    outs = []
    for varying in (False, True):
        o = po.OracleSystem(*shape)
        o.random_init(3, S0)
        o.sphere((4.0, 4.0, 3.0), 2.0, 0, 5.3, S0)
        o.set_params(A, B, C, 2.32, 0.5, 0.3, semantics=0)
        if varying:
            o.set_hfield_varying(np.tile(H, (n, 1)), 1.0, 1.0, 0.5)
        else:
            o.set_hfield(H, 1.0, 1.0, 0.5)
        outs.append((o.forces(), o.energy()[1]))
    np.testing.assert_allclose(outs[1][0], outs[0][0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(outs[1][1], outs[0][1], rtol=1e-15)
    h = np.tile(H, (n, 1))
    o = po.OracleSystem(*shape)
    o.random_init(3, S0)
    o.set_params(A, B, C, 2.32, semantics=1)
    o.set_hfield_varying(h, 1.0, 1.0, 0.5)
    f0 = o.forces()
    h[17] = (1.5, 0.2, -0.3)
    o.set_hfield_varying(h, 1.0, 1.0, 0.5)
    f1 = o.forces()
    changed = np.where(np.abs(f1 - f0).max(axis=1) > 0)[0]
    assert changed.tolist() == [17]


def _object_scene(phase):
    A, B, C, S0 = phase
    o = po.OracleSystem(10, 9, 8)
    o.random_init(77, S0)
    o.wall(2, 0, 1, 5.3, S0)
    o.sphere((5.0, 4.5, 4.0), 2.2, 0, 5.3, S0)
    o.sphere((1.5, 6.0, 5.5), 1.6, 1, 2.0, S0)
    return o


@live_ref
def test_object_sites_and_displacement_match_reference(phase):
    """cubicLattice::createBoundaryObject's surface sites (discovery order,
    duplicates dropped) and displaceBoundaryObject (types and Q) of the compiled
    reference, bit for bit, for moves along several axes and lengths."""
    A, B, C, S0 = phase
    o = _object_scene(phase)
    r = po.RefModel(10, 9, 8)
    for k in range(3):
        r.add_object(o.object_sites(k).astype(np.int32), 1 if k == 0 else 0, 5.3, S0)
    r.set_q(o.q())
    assert np.array_equal(r.types(), o.types())
    for k in range(3):
        assert np.array_equal(r.object_sites(k), o.object_sites(k))
        assert np.array_equal(r.object_sites(k, surface=True), o.object_sites(k, surface=True))
    for obj, direction, magnitude in ((1, 1, 2), (1, 4, 1), (2, 3, 3), (0, 5, 1), (1, 0, 1)):
        r.displace_object(obj, direction, magnitude)
        o.displace_object(obj, direction, magnitude)
        assert np.array_equal(r.types(), o.types()), (obj, direction, magnitude)
        assert np.array_equal(r.q(), o.q(), equal_nan=True), (obj, direction, magnitude)
        assert np.array_equal(r.object_sites(obj, surface=True), o.object_sites(obj, surface=True))


def test_object_force_properties(phase):
    """computeObjectForces on the oracle (the reference force class needs MPI and
    is not compiled here, so the stress formula is pinned by properties): a
    configuration with mirror symmetries has no net force; translating the whole
    scene leaves the force unchanged; a homeotropic sphere next to a planar
    wall, relaxed, is pushed along z only (the scene's x and y mirror symmetry)."""
    A, B, C, S0 = phase

    def scene(shift=(0, 0, 0), relax=0):
        o = po.OracleSystem(12, 12, 12)
        o.uniform_init((0.0, 0.0, 1.0), S0) if hasattr(o, "uniform_init") else o.director_init(
            np.tile([0.0, 0.0, 1.0, S0], (o.n, 1)))
        o.sphere((6.0 + shift[0], 6.0 + shift[1], 6.0 + shift[2]), 3.0, 0, 5.3, S0)
        o.set_params(A, B, C, 2.32, semantics=1)
        if relax:
            o.fire(max_iter=relax)
        return o

    f0 = scene().object_force(0)
    assert np.abs(f0).max() < 1e-12
    f1 = scene(relax=40).object_force(0)
    f2 = scene(shift=(2, 3, 1), relax=40).object_force(0)
    np.testing.assert_allclose(f2, f1, rtol=1e-10, atol=1e-12)
    assert np.abs(f1[:2]).max() < 1e-10
    # a planar wall below the sphere breaks the z mirror: a nonzero push along z
    o = po.OracleSystem(12, 12, 16)
    o.director_init(np.tile([0.0, 0.0, 1.0, S0], (o.n, 1)))
    o.wall(2, 0, 1, 5.3, S0)
    o.sphere((6.0, 6.0, 6.5), 3.0, 0, 5.3, S0)
    o.set_params(A, B, C, 2.32, semantics=1)
    o.fire(max_iter=60)
    f = o.object_force(1)
    assert np.abs(f[:2]).max() < 1e-9 * abs(f[2]) + 1e-13 and abs(f[2]) > 1e-6


def _q_matrix(q5):
    """Traceless symmetric Q from (Qxx, Qxy, Qxz, Qyy, Qyz)."""
    a, b, c, d, e = q5
    return np.array([[a, b, c], [b, d, e], [c, e, -a - d]])


def test_energy_density_sums_to_total_and_is_local():
    """The oracle's per-site energy density (landauDeGennesLC::energyDensity,
    filled by computeEnergyCPU, landauDeGennesLC.cpp:204-353): it sums to the
    total (pinned above against the reference's energy polynomials), is zero on
    object sites, equals a TrQ^2 + b TrQ^3 + c (TrQ^2)^2 computed here from the
    3x3 matrix on a uniform state (no gradients, no fields), and a change of Q at
    one site changes only that site's density and its six neighbours'."""
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    shape = (9, 8, 7)
    o = po.OracleSystem(*shape)
    o.random_init(5, S0)
    o.wall(2, 0, 1, 5.3, S0)
    o.sphere((4.0, 4.0, 3.5), 2.0, 0, 3.1, S0)
    o.set_params(A, B, C, 2.32, 1.0, 0.5, 0.2, 0.3, semantics=0)
    o.set_efield((0.0, 0.2, 0.5), 1.0, 1.0, 1.0)
    o.set_hfield((0.3, 0.0, -0.4), 1.0, 1.0, 0.5)
    dens = o.energy_density()
    total, _ = o.energy()
    assert dens.shape == (o.n,)
    np.testing.assert_allclose(dens.sum(), total, rtol=1e-12)
    objects = o.types() > 0
    assert objects.any() and np.all(dens[objects] == 0.0)
    assert np.all(dens[~objects] != 0.0)

    # locality: one site's Q changes its own density and its six neighbours'
    q = o.q()
    i = 7 + 9 * (1 + 8 * 5)       # (7, 1, 5): a bulk site away from the objects
    assert o.types()[i] == 0
    q2 = q.copy()
    q2[i] += np.array([0.01, -0.02, 0.03, 0.005, -0.01])
    changed = set(np.where(o.energy_density(q2) != dens)[0].tolist())
    nbrs = set(o.neighbors()[i].tolist())
    assert changed == {i} | nbrs

    # uniform state, no fields: the phase density alone, from the 3x3 matrix
    o2 = po.OracleSystem(*shape)
    n_dir = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    Qm = S0 * (np.outer(n_dir, n_dir) - np.eye(3) / 3.0)
    q5 = np.array([Qm[0, 0], Qm[0, 1], Qm[0, 2], Qm[1, 1], Qm[1, 2]])
    o2.set_q(np.tile(q5, (o2.n, 1)))
    o2.set_params(A, B, C, 2.32, 1.0, 0.5, 0.2, 0.3, semantics=0)
    M = _q_matrix(q5)
    t2, t3 = np.trace(M @ M), np.trace(M @ M @ M)
    want = 0.5 * A * t2 + B / 3.0 * t3 + 0.25 * C * t2 * t2
    np.testing.assert_allclose(o2.energy_density(), want, rtol=1e-13)


def _rotated_state(q, L, axis):
    """The periodic L^3 state q (site order, x fastest) rotated by 90 degrees about
    `axis`: site r goes to R r (mod L) and Q to R Q R^T.  Returns (q', R, T), with
    T the 5x5 map of the (Qxx, Qxy, Qxz, Qyy, Qyz) components."""
    c = [(1, 2), (2, 0), (0, 1)][axis]          # the rotated coordinate pair (a, b) -> (-b, a)
    R = np.eye(3)
    R[c[0], c[0]] = R[c[1], c[1]] = 0.0
    R[c[0], c[1]] = -1.0
    R[c[1], c[0]] = 1.0
    T = np.zeros((5, 5))
    for k in range(5):
        e = np.zeros(5)
        e[k] = 1.0
        M = R @ _q_matrix(e) @ R.T
        T[:, k] = [M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2]]
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    r = np.stack([x.ravel(), y.ravel(), z.ravel()])        # coordinates of sites 0..n-1
    r2 = (R.astype(int) @ r) % L
    dest = r2[0] + L * (r2[1] + L * r2[2])
    q2 = np.empty_like(q)
    q2[dest] = q @ T.T
    return q2, dest, T


@pytest.mark.parametrize("L4", [0.0, 0.2])
def test_energy_and_l1_l4_forces_are_covariant_under_lattice_rotations(L4):
    """Every distortion term of the reference is a rotational invariant (L4, the
    chiral one, under proper rotations), and its discrete energy uses central
    differences, which 90-degree lattice rotations map onto each other: on a
    periodic cube with a random Q field and no objects the energy with all of
    L1..L6 is unchanged by a rotation about each axis.  The L1 and L4 forces
    (with the phase force, metric corrected once as on the reference's GPU path)
    rotate with the field, F'(R r) = T F(r).  A transposed operand, a wrong axis
    or a sign in an energy polynomial or in those stencils breaks this.  (The
    reference's L2, L3 and L6 force stencils, which the oracle reproduces
    bit-exactly, are not covariant: they differ from a rotated field's by about
    1e-2 of the force scale, so they are not tested this way.)"""
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    L = 6
    o = po.OracleSystem(L, L, L)
    o.random_init(11, S0)
    o.set_params(A, B, C, 2.32, 1.0, 0.5, L4, 0.3, multi=True, semantics=1)
    q = o.q()
    e0, c0 = o.energy(q)
    for axis in range(3):
        q2, dest, T = _rotated_state(q, L, axis)
        e1, c1 = o.energy(q2)
        np.testing.assert_allclose(e1, e0, rtol=1e-12)
        np.testing.assert_allclose(c1, c0, rtol=1e-12, atol=1e-12 * abs(e0))
    o.set_params(A, B, C, 2.32, 0.0, 0.0, L4, 0.0, multi=True, semantics=1)
    f0 = o.forces(q)
    scale = np.abs(f0).max()
    for axis in range(3):
        q2, dest, T = _rotated_state(q, L, axis)
        f1 = o.forces(q2)
        np.testing.assert_allclose(f1[dest], f0 @ T.T, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_l1_plane_wave_closed_form(axis):
    """Closed form of the L1 force (lcForces.h:26-37 with the metric correction of
    the GPU path, landauDeGennesLC.cpp:137-160): with the phase term off, a plane
    wave Q(r) = eps sin(k r_axis) e on a periodic lattice (k = 2 pi m / L) has
    6Q - (sum of neighbours) = 2 (1 - cos k) Q; the shaping (s0 += s0 + s3,
    s3 += s0 + s3, the others doubled) followed by the metric (f0 <- 2/3 (2 f0 -
    f3), f3 <- 2/3 (2 f3 - f0)) is twice the identity, so F = -4 L1 (1 - cos k) Q
    exactly, along each axis."""
    shape = (8, 10, 12)
    L_axis = shape[axis]
    o = po.OracleSystem(*shape)
    L1 = 2.32
    o.set_params(0.0, 0.0, 0.0, L1, semantics=1)
    e = np.array([0.3, -0.2, 0.5, 0.1, -0.4])
    z, y, x = np.meshgrid(np.arange(shape[2]), np.arange(shape[1]), np.arange(shape[0]), indexing="ij")
    r = (x, y, z)[axis].ravel()
    for m in (1, 3):
        k = 2.0 * np.pi * m / L_axis
        q = 0.01 * np.outer(np.sin(k * r), e)
        f = o.forces(q)
        want = -4.0 * L1 * (1.0 - np.cos(k)) * q
        np.testing.assert_allclose(f, want, rtol=0, atol=1e-14 * np.abs(want).max())


REF_STRESS_SRC = "/root/reference/src/forces/landauDeGennesLCOtherForces.cpp"


@pytest.mark.skipif(not os.path.exists(REF_STRESS_SRC), reason="reference checkout absent")
def test_stress_tensor_matches_reference_expressions():
    """The oracle's one-constant stress tensor (computeObjectForces' per-site
    matrix) against the reference's own nine expressions
    (landauDeGennesLCOtherForces.cpp:214-251), read from its source text and
    evaluated on random derivative vectors, with the energy density added on the
    diagonal as there."""
    src = open(REF_STRESS_SRC).read()
    body = src[src.index("void landauDeGennesLC::computeStressTensors"):]
    call = body[body.index("stress.data[ii].set(") + len("stress.data[ii].set("):]
    depth, parts, cur = 0, [], ""
    for ch in call:
        if ch == "(":
            depth += 1
        elif ch == ")":
            if depth == 0:
                parts.append(cur)
                break
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
            continue
        cur += ch
    assert len(parts) == 9
    exprs = [compile(re.sub(r"firstDerivative\[(\d+)\]", r"d[\1]", p.strip()), "x%d" % k, "eval")
             for k, p in enumerate(parts)]
    lib = po.oracle_lib()
    rng = np.random.default_rng(9)
    for _ in range(50):
        d = np.ascontiguousarray(rng.standard_normal(15))
        L1 = float(rng.uniform(0.5, 3.0))
        e = float(rng.standard_normal())
        m = np.zeros(9)
        lib.orc_stress_tensor(po.dptr(d), L1, e, po.dptr(m))
        want = np.array([eval(x, {"L1": L1, "d": d}) for x in exprs])
        want[[0, 4, 8]] += e
        np.testing.assert_allclose(m, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_first_derivative_rules_on_a_quadratic_field():
    """computeFirstDerivatives (landauDeGennesLCOtherForces.cpp:15-145) on a field
    quadratic along z, Q(z) = z^2 e, with walls at z = 0, 4, 6 and 8: central
    differences in the bulk (z = 2); next to an object below (z = 1), the forward
    difference Q(z+1) - Q(z); next to an object above (z = 3), and with objects on
    both sides (z = 5, the upper one taking precedence), the backward difference
    Q(z) - Q(z-1); zero along x and y, in which the field is uniform.
    The values come from the field, so a wrong axis, side or sign fails."""
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    lx, ly, lz = 4, 3, 10
    o = po.OracleSystem(lx, ly, lz)
    for zw in (0, 4, 6, 8):
        o.wall(2, zw, 1, 0.0, S0)
    o.set_params(A, B, C, 2.32, semantics=1)
    e = np.array([0.3, -0.2, 0.5, 0.1, -0.4])
    z = np.repeat(np.arange(lz), lx * ly).astype(float)
    q = np.outer(z * z, e)
    types = o.types()
    lc = types <= 0
    q[~lc] = 0.0                                  # object sites keep their own Q; not read here
    d = o.first_derivatives(q).reshape(-1, 3, 5)
    Qz = lambda zz: zz * zz * e
    expect = {1: Qz(2) - Qz(1), 2: (Qz(3) - Qz(1)) / 2, 3: Qz(3) - Qz(2)}
    for zz in (1, 2, 3):
        i = int(np.where(z == zz)[0][0])
        assert types[i] <= 0
        np.testing.assert_allclose(d[i, 2], expect[zz], rtol=1e-15, atol=1e-15)
        np.testing.assert_array_equal(d[i, 0], 0.0)
        np.testing.assert_array_equal(d[i, 1], 0.0)
    # z = 5: objects at 4 and 6 -> the upper one wins: Q(5) - Q(4), with Q(4) the object's Q
    i5 = int(np.where(z == 5)[0][0])
    i4 = int(np.where(z == 4)[0][0])
    np.testing.assert_allclose(d[i5, 2], q[i5] - q[i4], rtol=1e-15, atol=1e-15)


def test_l2_l3_l6_stencils_are_not_covariant_on_smooth_fields():
    """Register of a reference property the oracle and the GPU path reproduce:
    on a smooth periodic field (a few Fourier modes of wavelength 24 sites) the
    L1 and L4 distortion forces rotate with the field under 90-degree lattice
    rotations, while the reference's L2, L3 and L6 stencils
    (src/forces/lcForces.h:66-1043, bit-exact in the oracle) miss by 10-35% of
    their force scale (20-32% on this field), so they are not a discretisation of the gradient of the
    rotation-invariant energy the reference reports."""
    L = 24
    o = po.OracleSystem(L, L, L)
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    rng = np.random.default_rng(2)
    q = np.zeros((L ** 3, 5))
    for c in range(5):
        f = np.zeros((L, L, L))
        for _ in range(3):
            k = rng.integers(-1, 2, size=3)
            f += rng.normal() * 0.1 * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) / L + rng.uniform(0, 6.3))
        q[:, c] = f.ravel()
    mismatch = {}
    for name, pars in (("L1", (1, 0, 0, 0, 0)), ("L2", (0, 1, 0, 0, 0)), ("L3", (0, 0, 1, 0, 0)),
                       ("L4", (0, 0, 0, 1, 0)), ("L6", (0, 0, 0, 0, 1))):
        o.set_params(0.0, 0.0, 0.0, *pars, multi=True, semantics=1)
        f0 = o.forces(q)
        worst = 0.0
        for axis in range(3):
            q2, dest, T = _rotated_state(q, L, axis)
            worst = max(worst, np.abs(o.forces(q2)[dest] - f0 @ T.T).max() / np.abs(f0).max())
        mismatch[name] = worst
    assert mismatch["L1"] < 1e-12 and mismatch["L4"] < 1e-12, mismatch
    for name in ("L2", "L3", "L6"):
        assert 0.05 < mismatch[name] < 0.5, mismatch


@live_ref
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_defect_measures_match_reference(phase, kind):
    """computeDefectMeasures (qTensorLatticeModel.cpp:95-133) of the compiled
    reference model for all three kinds (largest eigenvalue, determinant,
    degree of biaxiality) on a random biaxial field (the biaxiality measure is
    zero on the uniaxial seeded start): the oracle's equal them on every site."""
    shape = (7, 6, 5)
    n = shape[0] * shape[1] * shape[2]
    q = np.random.default_rng(5).normal(0.0, 0.3, size=(n, 5))
    r = po.RefModel(*shape)
    r.set_q(q)
    o = po.OracleSystem(*shape)
    o.set_q(q)
    want = r.defects(kind)
    assert np.abs(want).max() > 0
    np.testing.assert_allclose(o.defects(kind), want, rtol=1e-12, atol=1e-15 * np.abs(want).max())
