"""The FIRE pass that keeps only Q and v between iterations (tiles 4-6,
fire_stream.hip) against the CPU oracle: it recomputes each bulk site's force
from Q instead of reading it, so its trajectories must follow the oracle's as
the fused pass's do.  Tolerances as in test_gpu_parity.py: traced forceMax
within 1e-10 relative, Q and v within 1e-11 in FP64; FP32 traces within 1e-4
of the FP64 oracle's.  Odd lattice sizes exercise partial tiles; the force and
velocity left in the lattice after a run are the reference's (F of the final
state, velocity mixed)."""
import numpy as np
import pytest

import qmin
from test_gpu_parity import A, B, C, MULTI, ONE, S0, build_pair

pytestmark = pytest.mark.gpu

L4 = (2.32, 1.0, 0.5, 0.2, 0.0)
ALL = (2.32, 1.0, 0.5, 0.2, 0.3)


@pytest.mark.parametrize("tile", [4, 5])
@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("params", [ONE, MULTI, L4], ids=["L1", "L1-L3", "L1-L4"])
def test_stream_fire_trajectory_fp64(tile, semantics, params):
    print("stream FIRE fp64: tile %d semantics %d params %s" % (tile, semantics, params), flush=True)
    o, g, f = build_pair((37, 21, 13), params, semantics, True)
    g.set_tile(tile)
    it_ref, tr_ref = o.fire(max_iter=40)
    m = qmin.FIRE(max_iter=40)
    m.set_trace(64)
    st = m.minimize(f, g)
    assert st.iterations == it_ref == 40
    tr = m.trace()
    np.testing.assert_array_equal(tr[:, 5], tr_ref[:, 5])
    np.testing.assert_allclose(tr[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(tr[:, 2:4], tr_ref[:, 2:4], rtol=1e-12)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(g.v(), o.v(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(g.f(), o.f(), rtol=0, atol=1e-11)


def test_stream_fire_fields_and_varying_field():
    print("stream FIRE with E field and a varying H field", flush=True)
    o, g, f = build_pair((34, 19, 12), L4, 0, True, efield=(0.1, -0.2, 0.5))
    h = np.random.default_rng(3).normal(0.0, 0.4, size=(g.n, 3))
    o.set_hfield_varying(h, 1.0, 1.0, 0.6)
    g.set_hfield(h)
    f.set_varying_hfield(1.0, 1.0, 0.6)
    g.set_tile(4)
    it_ref, tr_ref = o.fire(max_iter=30)
    m = qmin.FIRE(max_iter=30)
    m.set_trace(32)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("tile", [4, 5, 6])
def test_stream_fire_fp32_tracks_fp64(tile):
    print("stream FIRE fp32: tile %d" % tile, flush=True)
    o, g, f = build_pair((70, 19, 11), MULTI, 1, True, precision=qmin.QM_FP32)
    g.set_tile(tile)
    it_ref, tr_ref = o.fire(max_iter=30)
    m = qmin.FIRE(max_iter=30)
    m.set_trace(32)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-4)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-5)


def test_stream_fire_continues_and_moves_objects():
    """Two minimize calls equal one, and minimize - move an object - minimize
    follows the oracle: the state between calls is the reference's."""
    print("stream FIRE continuation and object move", flush=True)
    o, g, f = build_pair((36, 20, 14), MULTI, 0, True)
    g.set_tile(4)
    o.fire(max_iter=40)
    m = qmin.FIRE(max_iter=20)
    m.minimize(f, g)
    m.set_max_iterations(40)
    m.minimize(f, g)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    g.displace_object(2, 3, 1)
    o.displace_object(2, 3, 1)
    qmin.FIRE(max_iter=20).minimize(f, g)
    o.fire(max_iter=20)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(g.v(), o.v(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("tile", [4, 5])
def test_stream_run_iterations_stores_the_final_force(tile):
    """run_iterations has no tolerance, so its last pass is known and stores F
    itself (no separate force pass at the end): Q, the mixed velocity and F left
    in the lattice are the oracle's after the same iterations, after a second
    call too."""
    print("stream run_iterations: final force, tile %d" % tile, flush=True)
    _, g, f = build_pair((37, 21, 13), MULTI, 1, True)
    g.set_tile(tile)
    m = qmin.FIRE(max_iter=10 ** 6)
    done = 0
    for n in (17, 9):
        st = m.run_iterations(f, g, n)
        done += n
        assert not st.nanDetected and st.iterations == done
        o, _, _ = build_pair((37, 21, 13), MULTI, 1, True)   # the oracle runs from the start in one call
        o.fire(max_iter=done, fc=0.0)
        np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
        np.testing.assert_allclose(g.v(), o.v(), rtol=0, atol=1e-11)
        np.testing.assert_allclose(g.f(), o.f(), rtol=0, atol=1e-11)


def test_stream_tile_with_l6_uses_the_fused_pass():
    """The streaming pass has no L6 form: with L6 a streaming tile setting
    leaves FIRE on the fused pass, which the oracle still matches."""
    o, g, f = build_pair((24, 18, 12), ALL, 1, True)
    g.set_tile(4)
    it_ref, tr_ref = o.fire(max_iter=20)
    m = qmin.FIRE(max_iter=20)
    m.set_trace(32)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


def test_stream_and_fused_pass_agree_at_256():
    """At the benchmark's size (256^3, walls and sphere, L1-L3) both passes run 30
    iterations from the same start: forceMax traces agree to 1e-12, Q to 1e-13."""
    print("stream vs fused pass at 256^3", flush=True)
    import bench
    out = []
    for tile in (-1, 4):
        lat, force = bench.build_workload(qmin, 256, qmin.QM_FP64)
        lat.set_tile(tile)
        m = qmin.FIRE(fc=1e-12, max_iter=30, **bench.fire_params())
        m.set_trace(32)
        m.minimize(force, lat)
        out.append((m.trace()[:, 0], lat.q()))
        del lat
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=1e-12)
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=0, atol=1e-13)
