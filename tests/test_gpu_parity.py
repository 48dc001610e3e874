"""HIP path (libqmin.so on the MI355X) against the CPU oracle on seeded inputs.

Tolerances: FP64 forces within 1e-12 relative at every site (assert_forces_close;
FP32 states within 2e-5 of the lattice's largest force); FP64 energies within 1e-12
relative; FIRE trajectories within 1e-10 relative per traced scalar (FP64
reduction order differs from the oracle's serial sums); integer data bit-exact.
"""
import numpy as np
import pytest

import pyoracle as po
import qmin

pytestmark = pytest.mark.gpu

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)

ONE = (2.32, 0.0, 0.0, 0.0, 0.0)
MULTI = (2.32, 0.5, 0.3, 0.0, 0.0)
ALL = (2.32, 1.0, 0.5, 0.2, 0.3)


def build_pair(shape, params, semantics, objects=True, efield=None, seed=1234, precision=qmin.QM_FP64):
    o = po.OracleSystem(*shape)
    o.random_init(seed, S0)
    g = qmin.Lattice(*shape, precision=precision)
    g.random_init(seed, S0)
    if objects:
        lx, ly, lz = shape
        o.wall(2, 0, 1, 5.3, S0)
        g.wall(2, 0, 1, 5.3, S0)
        o.wall(2, lz - 1, 1, 5.3, S0)
        g.wall(2, lz - 1, 1, 5.3, S0)
        c = (lx / 2.0, ly / 2.0, lz / 2.0)
        r = min(lx, ly, lz) / 4.0
        o.sphere(c, r, 0, 5.3, S0)
        g.sphere(c, r, 0, 5.3, S0)
    o.set_params(A, B, C, *params, semantics=semantics)
    f = qmin.LandauDeGennesLC(A, B, C, *params, semantics=semantics)
    if efield is not None:
        o.set_efield(efield, 1.0, 1.0, 1.0)
        f.set_efield(efield, 1.0, 1.0, 1.0)
    return o, g, f


def assert_forces_close(got, ref, rel=1e-12):
    """FP64 forces against the oracle, per site: |F_got - F_ref| <= rel |F_ref|
    (2-norms of the five components) wherever F_ref is not zero, and F_got == 0
    exactly where it is (object sites).  Measured on the seeded test lattices:
    at most 5.4e-15."""
    got = np.asarray(got).reshape(-1, 5)
    ref = np.asarray(ref).reshape(-1, 5)
    nref = np.linalg.norm(ref, axis=1)
    err = np.linalg.norm(got - ref, axis=1)
    zero = nref == 0
    assert np.all(err[zero] == 0), "force on %d zero-force sites" % int((err[zero] != 0).sum())
    rel_err = err[~zero] / nref[~zero]
    worst = int(np.argmax(rel_err)) if rel_err.size else 0
    assert rel_err.size == 0 or rel_err[worst] <= rel, \
        "per-site relative error %g > %g (|F_ref| there %g)" % (rel_err[worst], rel, nref[~zero][worst])


def assert_fp32_forces_close(got, ref, rel=2e-5):
    """FP32 state against the FP64 oracle: the state itself is rounded to float,
    so the bound is relative to the largest force of the lattice, not per site
    (where a small force is the difference of terms whose rounding is larger)."""
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max()
    assert err <= rel * scale, "max |dF| = %g, scale %g" % (err, scale)


def test_lattice_setup_bit_exact():
    shape = (13, 9, 11)
    o, g, _ = build_pair(shape, MULTI, 0)
    assert np.array_equal(g.types(), o.types())
    assert np.array_equal(g.neighbors(), o.neighbors())
    assert np.array_equal(g.q(), o.q()), "seeded initial Q must be bit-exact"
    assert g.n_active() == o.n_active()


@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("params,objects,shape", [
    (ONE, False, (12, 10, 9)),
    (ONE, True, (35, 17, 12)),
    (MULTI, False, (33, 9, 7)),
    (MULTI, True, (40, 20, 16)),
    (ALL, True, (24, 18, 14)),
])
def test_forces_fp64(semantics, params, objects, shape):
    o, g, f = build_pair(shape, params, semantics, objects)
    ref = o.forces()
    got = f.compute_forces(g)
    assert_forces_close(got, ref)


def _boundary_cases(types, shape):
    """getBoundaryCase of every type -1 site (lcForces.h; xx, yy, zz = 1 for an
    object above, 2 for one below, below winning), from the site types."""
    lx, ly, lz = shape
    obj = types.reshape(lz, ly, lx) > 0
    t = types.reshape(lz, ly, lx)
    code = np.zeros(t.shape, dtype=int)
    for axis, w in ((2, 1), (1, 4), (0, 16)):
        up = np.roll(obj, -1, axis=axis)
        down = np.roll(obj, 1, axis=axis)
        code += w * np.where(down, 2, np.where(up, 1, 0))
    return set(np.unique(code[t == -1]).tolist())


@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("params", [MULTI, ALL], ids=["L1-L3", "L1-L6"])
def test_forces_cover_every_boundary_case(semantics, params):
    """Two spheres (radius 5.6 homeotropic, 6.7 degenerate planar) whose surface
    sites take all 26 object-neighbour cases of the boundary stencils, and a
    one-site object; bulk sites near them need one-sided neighbour derivatives
    (the boundary pass's case 0; 519 here): per-site forces against the oracle,
    and FIRE iterations."""
    shape = (36, 26, 24)
    o = po.OracleSystem(*shape)
    g = qmin.Lattice(*shape)
    for sys_ in (o, g):
        sys_.random_init(77, S0)
        sys_.sphere((10.3, 12.6, 11.8), 5.6, 0, 5.3, S0)
        sys_.sphere((26.6, 13.2, 12.4), 6.7, 1, 3.1, S0)
        sys_.sphere((18.5, 4.5, 20.5), 0.9, 0, 8.0, S0)
    types = o.types()
    assert np.array_equal(g.types(), types)
    cases = _boundary_cases(types, shape)
    assert cases == set(range(64)) - {0} - {c for c in range(64) if c % 4 == 3 or (c // 4) % 4 == 3 or c // 16 >= 3}
    assert len(cases) == 26
    # bulk sites next to a boundary site that has an object beside it (across
    # the bulk site's line to it): their neighbour derivatives are one-sided
    t3 = types.reshape(shape[2], shape[1], shape[0])
    obj = t3 > 0
    one_sided = np.zeros(t3.shape, bool)
    for ax in range(3):
        for sgn in (1, -1):
            nb_is_surface = np.roll(t3, -sgn, axis=ax) == -1
            beside = np.zeros(t3.shape, bool)
            for ax2 in range(3):
                if ax2 != ax:
                    for s2 in (1, -1):
                        beside |= np.roll(np.roll(obj, -sgn, axis=ax), -s2, axis=ax2)
            one_sided |= (t3 == 0) & nb_is_surface & beside
    assert one_sided.sum() >= 12
    o.set_params(A, B, C, *params, semantics=semantics)
    f = qmin.LandauDeGennesLC(A, B, C, *params, semantics=semantics)
    assert_forces_close(f.compute_forces(g), o.forces())
    it_ref, tr_ref = o.fire(max_iter=15)
    m = qmin.FIRE(max_iter=15)
    m.set_trace(16)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


def test_forces_with_field():
    o, g, f = build_pair((20, 16, 18), ALL, 0, True, efield=(0.0, 0.0, 0.5))
    assert_forces_close(f.compute_forces(g), o.forces())
    o, g, f = build_pair((20, 16, 18), MULTI, 1, True, efield=(0.3, -0.2, 0.5))
    assert_forces_close(f.compute_forces(g), o.forces())


def test_forces_fp32_close():
    o, g, f = build_pair((40, 20, 16), MULTI, 1, True, precision=qmin.QM_FP32)
    ref = o.forces()
    got = f.compute_forces(g)
    assert_fp32_forces_close(got, ref)


@pytest.mark.parametrize("params", [ONE, MULTI, ALL])
def test_energy_fp64(params):
    o, g, f = build_pair((24, 18, 14), params, 0, True, efield=(0.0, 0.1, 0.4))
    e_ref, c_ref = o.energy()
    e_got, c_got = f.compute_energy(g)
    np.testing.assert_allclose(e_got, e_ref, rtol=1e-12)
    np.testing.assert_allclose(c_got, c_ref, rtol=1e-11, atol=1e-12 * abs(e_ref))


@pytest.mark.parametrize("params", [ONE, MULTI, ALL])
def test_energy_density_fp64(params):
    """landauDeGennesLC::energyDensity: every site's energy against the oracle's
    (pinned there by its sum and locality), zero on object sites, summing to
    compute_energy's total."""
    o, g, f = build_pair((24, 18, 14), params, 0, True, efield=(0.0, 0.1, 0.4))
    want = o.energy_density()
    got = f.energy_density(g)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert np.all(got[o.types() > 0] == 0.0)
    total, _ = f.compute_energy(g)
    np.testing.assert_allclose(got.sum(), total, rtol=1e-12)


def test_energy_density_fp32_close():
    o, g, f = build_pair((24, 18, 14), ALL, 0, True, efield=(0.0, 0.1, 0.4), precision=qmin.QM_FP32)
    want = o.energy_density()
    got = f.energy_density(g)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.abs(want).max())


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_defects(kind):
    """computeDefectMeasures (qTensorLatticeModel.cpp:95-133): 0 the largest
    eigenvalue, 1 the determinant, 2 the degree of biaxiality, on the
    liquid-crystal sites of a lattice with objects and the same random biaxial
    field on both sides (the biaxiality measure is zero on the uniaxial seeded
    start)."""
    o, g, _ = build_pair((16, 12, 10), MULTI, 0, True)
    q = np.random.default_rng(8).normal(0.0, 0.3, size=(g.n, 5))
    o.set_q(q)
    g.set_q(q)
    lc = o.types() <= 0
    want = o.defects(kind)[lc]
    assert np.abs(want).max() > 0
    np.testing.assert_allclose(g.defects(kind)[lc], want, rtol=1e-12, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("params", [ONE, MULTI, ALL])
def test_fire_trajectory_fp64(semantics, params):
    shape = (20, 18, 12)
    o, g, f = build_pair(shape, params, semantics, True)
    it_ref, tr_ref = o.fire(max_iter=40)
    m = qmin.FIRE(max_iter=40)
    m.set_trace(64)
    st = m.minimize(f, g)
    assert st.iterations == it_ref == 40
    tr = m.trace()
    assert tr.shape == tr_ref.shape
    np.testing.assert_array_equal(tr[:, 5], tr_ref[:, 5])       # N+ counter
    np.testing.assert_allclose(tr[:, 0], tr_ref[:, 0], rtol=1e-10)  # forceMax
    np.testing.assert_allclose(tr[:, 2:4], tr_ref[:, 2:4], rtol=1e-12)  # alpha, dt
    np.testing.assert_allclose(tr[:, 1], tr_ref[:, 1], rtol=1e-9, atol=1e-12 * np.abs(tr_ref[:, 1]).max())
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(g.v(), o.v(), rtol=0, atol=1e-11)


def test_fire_one_constant_periodic_50():
    """Config 1 of the workload set: 50^3 periodic, random start seed 1234, 100 FIRE steps."""
    o, g, f = build_pair((50, 50, 50), ONE, 0, False)
    it_ref, tr_ref = o.fire(max_iter=100)
    m = qmin.FIRE(max_iter=100)
    m.set_trace(128)
    m.minimize(f, g)
    tr = m.trace()
    np.testing.assert_allclose(tr[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    e_ref, _ = o.energy()
    e_got, _ = f.compute_energy(g)
    np.testing.assert_allclose(e_got, e_ref, rtol=1e-12)


def test_fire_stops_on_tolerance_like_reference():
    o, g, f = build_pair((16, 12, 10), ONE, 0, True)
    it_ref, tr_ref = o.fire(fc=0.02, max_iter=5000)
    m = qmin.FIRE(fc=0.02, max_iter=5000)
    st = m.minimize(f, g)
    assert st.iterations == it_ref
    assert st.converged and st.forceMax <= 0.02
    np.testing.assert_allclose(st.forceMax, tr_ref[-1, 0], rtol=1e-10)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_converged_state_matches_oracle(mode):
    """At a fixed tolerance (forceMax < 1e-11, walls + sphere, L1-L3) the converged
    energy and the largest-eigenvalue defect field against the oracle's FIRE run.

    FP64 follows the oracle's trajectory and stops at the same iteration (2856):
    energy and defect field within 1e-9 relative (measured: 1.6e-15 and 1.3e-15).
    Mixed runs FP32 to forceMax < 1e-7 (on 4320 sites FP32 rounding alone leaves
    forceMax = |F|/N near 1e-8, so the switch comes earlier than at 256^3), then
    converts, which gives the object sites back their FP64 boundary values, and
    polishes in FP64 to 1e-11.  It ends in the same minimum at another point
    within the tolerance: energy within 1e-9 relative (measured 6.3e-11; 7.0e-9
    while the converted state kept FP32-rounded boundary values), defect field
    1.2e-8 of its largest value, bound 1e-7."""
    fc = 1e-11
    shape = (20, 18, 12)
    prec = qmin.QM_FP64 if mode == "fp64" else qmin.QM_FP32
    o, g, f = build_pair(shape, MULTI, 1, True, precision=prec)
    it_ref, tr_ref = o.fire(fc=fc, max_iter=200000)
    assert tr_ref[-1, 0] < fc
    m = qmin.FIRE(fc=fc if mode == "fp64" else 1e-7, max_iter=200000)
    st = m.minimize(f, g)
    if mode == "mixed":
        assert st.converged
        g.set_precision(qmin.QM_FP64)
        m.set_force_cutoff(fc)
        st = m.minimize(f, g)
    assert st.converged and st.forceMax < fc
    if mode == "fp64":
        assert st.iterations == it_ref
    e_ref, _ = o.energy()
    e_got, _ = f.compute_energy(g)
    lc = o.types() <= 0
    d_ref, d_got = o.defects(0)[lc], g.defects(0)[lc]
    d_tol = 1e-9 if mode == "fp64" else 1e-7
    assert abs(e_got - e_ref) <= 1e-9 * abs(e_ref)
    assert np.abs(d_got - d_ref).max() <= d_tol * np.abs(d_ref).max()


def test_random_uniform_and_function_inits_bit_exact():
    """openQmin -z 1 and -z 3 initial states on a lattice with objects: only the
    liquid-crystal sites change, bit-exact with the oracle; a callable director
    function gives the same as its values."""
    from conftest import example_directors
    shape = (20, 18, 12)
    o, g, f = build_pair(shape, MULTI, 1, True)
    o.random_uniform_init(4321, S0)
    g.random_uniform_init(4321, S0)
    assert np.array_equal(g.q(), o.q())
    n4 = example_directors(shape)
    o.director_init(n4)
    g.director_init(n4)
    assert np.array_equal(g.q(), o.q())
    lx, ly, lz = shape
    fn = np.array([[x, y + 1.0, 1.0, 0.5] for z in range(lz) for y in range(ly) for x in range(lx)])
    g.director_init(lambda x, y, z: (x, y + 1.0, 1.0, 0.5))
    o.director_init(fn)
    assert np.array_equal(g.q(), o.q())


def test_object_sites_displacement_and_force_match_oracle():
    """Object and surface site lists, displaceBoundaryObject (types and Q, bit
    for bit) and the stress-tensor force on an object (computeObjectForces) on
    the device, against the oracle; multi-constant models are refused, as the
    reference computes the stress only for one constant."""
    shape = (16, 14, 12)
    o, g, f = build_pair(shape, ONE, 1, True)
    for k in range(3):
        assert np.array_equal(g.object_sites(k), o.object_sites(k))
        assert np.array_equal(g.object_sites(k, surface=True), o.object_sites(k, surface=True))
    for k in range(3):
        np.testing.assert_allclose(f.object_force(g, k), o.object_force(k), rtol=1e-12, atol=1e-12)
    for obj, direction, magnitude in ((2, 1, 2), (2, 4, 1), (1, 5, 1)):
        g.displace_object(obj, direction, magnitude)
        o.displace_object(obj, direction, magnitude)
        assert np.array_equal(g.types(), o.types())
        assert np.array_equal(g.q(), o.q(), equal_nan=True)
        assert np.array_equal(g.object_sites(obj, surface=True), o.object_sites(obj, surface=True))
    np.testing.assert_allclose(f.object_force(g, 2), o.object_force(2), rtol=1e-12, atol=1e-12)
    # after FIRE the force follows the oracle's trajectory
    m = qmin.FIRE(max_iter=20)
    m.minimize(f, g)
    o.fire(max_iter=20)
    np.testing.assert_allclose(f.object_force(g, 2), o.object_force(2), rtol=1e-9, atol=1e-11)
    multi = qmin.LandauDeGennesLC(A, B, C, *MULTI, semantics=1)
    with pytest.raises(qmin.QminError):
        multi.object_force(g, 2)


@pytest.mark.parametrize("semantics,params", [(1, ONE), (0, MULTI)])
def test_fire_displace_fire_matches_oracle(semantics, params):
    """The moving-colloid workflow: minimize, move an object, minimize again.
    energyMinimizerFIRE mixes the velocity with the force at the end of each step
    (energyMinimizerFIRE.cpp:113-116), so the velocity that the second run starts
    from holds the force from before the move, not after it."""
    o, g, f = build_pair((18, 16, 14), params, semantics, True)
    qmin.FIRE(max_iter=20).minimize(f, g)
    o.fire(max_iter=20)
    # (no read-back of v here: reading it would make the mix concrete itself)
    g.displace_object(2, 1, 1)
    o.displace_object(2, 1, 1)
    assert np.array_equal(g.types(), o.types())
    # a force evaluation in between changes nothing either
    f.compute_forces(g)
    m = qmin.FIRE(max_iter=20)
    m.set_trace(32)
    m.minimize(f, g)
    it_ref, tr_ref = o.fire(max_iter=20)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(g.v(), o.v(), rtol=0, atol=1e-11)


def test_fire_continues_across_calls():
    """Two minimize calls of 20 iterations equal one of 40 (state persists on the device)."""
    o, g, f = build_pair((18, 14, 12), MULTI, 0, True)
    o.fire(max_iter=40)
    m = qmin.FIRE(max_iter=20)
    m.minimize(f, g)
    m.set_max_iterations(40)
    m.minimize(f, g)
    assert m.state().iterations == 40
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


def test_precision_conversion_round_trip():
    """FP64 -> FP32 rounds Q, v, F to float; FP32 -> FP64 keeps the float values
    of the liquid crystal and gives the object sites back their FP64 Q (the
    fixed boundary values); FIRE continues."""
    o, g, f = build_pair((14, 12, 10), MULTI, 1, True)
    m = qmin.FIRE(max_iter=5)
    m.minimize(f, g)
    q64, v64, f64 = g.q(), g.v(), g.f()
    g.set_precision(qmin.QM_FP32)
    assert np.array_equal(g.q(), q64.astype(np.float32).astype(np.float64))
    assert np.array_equal(g.v(), v64.astype(np.float32).astype(np.float64))
    assert np.array_equal(g.f(), f64.astype(np.float32).astype(np.float64))
    g.set_precision(qmin.QM_FP64)
    obj = g.types() > 0
    q = g.q()
    assert obj.any() and not np.array_equal(q64[obj], q64[obj].astype(np.float32).astype(np.float64))
    assert np.array_equal(q[obj], q64[obj])
    assert np.array_equal(q[~obj], q64[~obj].astype(np.float32).astype(np.float64))
    m.set_max_iterations(10)
    st = m.minimize(f, g)
    assert st.iterations == 10 and np.isfinite(st.forceMax)


def test_fp32_lattice_converted_to_fp64_has_exact_object_q():
    """Objects added to an FP32 lattice, a seeded start (which downloads Q and
    uploads it again), FIRE in FP32, then the conversion: the object sites hold
    the FP64 boundary values bit for bit, as on a lattice built in FP64, so the
    FP64 polish minimizes the same energy as an FP64 run."""
    o, g, f = build_pair((18, 14, 12), MULTI, 1, True)
    _, g32, f32 = build_pair((18, 14, 12), MULTI, 1, True, precision=qmin.QM_FP32)
    qmin.FIRE(max_iter=20).minimize(f32, g32)
    g32.set_precision(qmin.QM_FP64)
    obj = g.types() > 0
    assert np.array_equal(g32.types(), g.types())
    assert np.array_equal(g32.q()[obj], o.q()[obj])


@pytest.mark.parametrize("params", [MULTI, ALL], ids=["L1-L3", "L1-L6"])
@pytest.mark.parametrize("precision", [qmin.QM_FP64, qmin.QM_FP32])
def test_tile_shapes_agree(precision, params):
    """Every compiled tile shape of the fused pass (32x8, 64x4, 32x16, 64x8) gives the
    same per-site forces (bit for bit: the per-site arithmetic does not depend on
    the tiling) and FIRE traces equal to reduction rounding; odd sizes exercise
    partial tiles.  With L6 the 512-thread tiles are never the default, but the
    tile setter and the autotuner can choose them."""
    shape = (70, 19, 9)
    o, g, f = build_pair(shape, params, 1, True, precision=precision)
    got = []
    for tile in (0, 1, 2, 3):
        g.set_tile(tile)
        got.append(f.compute_forces(g))
    for other in got[1:]:
        assert np.array_equal(got[0], other)
    if precision == qmin.QM_FP64:
        assert_forces_close(got[0], o.forces())
    traces = []
    for tile in (0, 1, 2, 3):
        o2, g2, f2 = build_pair(shape, params, 1, True, precision=precision)
        g2.set_tile(tile)
        m = qmin.FIRE(max_iter=30)
        m.set_trace(32)
        m.minimize(f2, g2)
        traces.append(m.trace())
    for t in traces[1:]:
        np.testing.assert_allclose(t[:, 0], traces[0][:, 0], rtol=1e-12 if precision == qmin.QM_FP64 else 1e-6)


def test_autotune_leaves_state_and_results_unchanged():
    o, g, f = build_pair((40, 24, 20), MULTI, 1, True)
    m = qmin.FIRE(max_iter=10)
    m.minimize(f, g)
    o.fire(max_iter=10)
    q0, v0 = g.q(), g.v()
    tile, zchunk, ms = g.autotune(f)
    assert tile in (0, 1, 2) and zchunk in (0, 8, 16, 32, 64) and ms > 0
    assert np.array_equal(g.q(), q0) and np.array_equal(g.v(), v0)
    # the run continues on the tuned shape exactly as the oracle's uninterrupted one
    m.set_max_iterations(25)
    m.minimize(f, g)
    o2, _, _ = build_pair((40, 24, 20), MULTI, 1, True)
    it_ref, tr_ref = o2.fire(max_iter=25)
    np.testing.assert_allclose(g.q(), o2.q(), rtol=0, atol=1e-11)


def _random_field(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.4, size=(n, 3))


@pytest.mark.parametrize("semantics,params,precision", [(0, ALL, qmin.QM_FP64), (1, MULTI, qmin.QM_FP64),
                                                        (1, ONE, qmin.QM_FP32)])
def test_spatially_varying_field_forces_energy(semantics, params, precision):
    shape = (26, 18, 14)
    o, g, f = build_pair(shape, params, semantics, True, precision=precision)
    h = _random_field(g.n, 5)
    o.set_hfield_varying(h, 1.0, 1.0, 0.7)
    g.set_hfield(h)
    f.set_varying_hfield(1.0, 1.0, 0.7)
    if precision == qmin.QM_FP64:
        assert_forces_close(f.compute_forces(g), o.forces())
    else:
        assert_fp32_forces_close(f.compute_forces(g), o.forces())
    e_ref, c_ref = o.energy()
    e_got, c_got = f.compute_energy(g)
    np.testing.assert_allclose(e_got, e_ref, rtol=1e-12 if precision == qmin.QM_FP64 else 1e-5)
    np.testing.assert_allclose(c_got[4], c_ref[4], rtol=1e-11 if precision == qmin.QM_FP64 else 1e-5)


def test_spatially_varying_field_fire_and_file(tmp_path):
    shape = (20, 16, 12)
    o, g, f = build_pair(shape, MULTI, 0, True)
    h = _random_field(g.n, 9)
    path = tmp_path / "field_x0y0z0.txt"
    with open(path, "w") as fh:
        for i in range(g.n):
            x, y, z = i % 20, (i // 20) % 16, i // (20 * 16)
            fh.write("%d %d %d %.17g %.17g %.17g\n" % (x, y, z, h[i, 0], h[i, 1], h[i, 2]))
    o.set_hfield_varying(h, 1.0, 1.0, 0.5)
    g.load_hfield(path)
    f.set_varying_hfield(1.0, 1.0, 0.5)
    it_ref, tr_ref = o.fire(max_iter=30)
    m = qmin.FIRE(max_iter=30)
    m.set_trace(32)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("kind", ["adam", "nesterov"])
def test_spatially_varying_field_adam_nesterov(kind):
    """The Adam and Nesterov passes (force and update in one pass) with a
    spatially varying H field and a uniform E field, L1-L6, objects: traces and
    final Q against the oracle."""
    shape = (22, 16, 12)
    o, g, f = build_pair(shape, ALL, 0, True, efield=(0.1, 0.0, 0.4))
    h = _random_field(g.n, 13)
    o.set_hfield_varying(h, 1.0, 1.0, 0.6)
    g.set_hfield(h)
    f.set_varying_hfield(1.0, 1.0, 0.6)
    if kind == "adam":
        _, tr_ref = o.adam(dt=1e-3, max_iter=25)
        m = qmin.Adam(dt=1e-3, max_iter=25)
    else:
        _, tr_ref = o.nesterov(dt=1e-2, mu=0.5, scheduled=True, max_iter=25)
        m = qmin.Nesterov(dt=1e-2, mu=0.5, scheduled=True, max_iter=25)
    m.set_trace(32)
    it, _ = m.minimize(f, g)
    assert it == 25
    np.testing.assert_allclose(m.trace(), tr_ref, rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


def test_varying_field_requires_lattice_field():
    _, g, f = build_pair((12, 10, 9), ONE, 1, False)
    f.set_varying_hfield(1.0, 1.0, 0.5)
    with pytest.raises(qmin.QminError):
        f.compute_forces(g)
