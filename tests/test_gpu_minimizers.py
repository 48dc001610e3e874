"""Adam, Nesterov AG and online L-BFGS on the MI355X against the CPU oracle.

The oracle's minimizers reproduce the reference's traces bit for bit
(test_oracle_reference.py); here the device runs the same iterations from the
same seeded state.  Tolerances: per-iteration forceMax within 1e-10 relative and
final Q within 1e-11 absolute in FP64 (the device sums its reductions in a tree
order, the oracle serially); FP32 state within 1e-4 relative on forceMax.
"""
import numpy as np
import pytest

import qmin
from test_gpu_parity import ALL, MULTI, ONE, assert_forces_close, build_pair

pytestmark = pytest.mark.gpu


def check(tr, tr_ref, g, o, rtol=1e-10, atol_q=1e-11):
    assert tr.shape == tr_ref.shape, (tr.shape, tr_ref.shape)
    np.testing.assert_allclose(tr, tr_ref, rtol=rtol)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=atol_q)


@pytest.mark.parametrize("params", [ONE, MULTI, ALL])
def test_adam_trajectory(params):
    o, g, f = build_pair((20, 18, 12), params, 0, True)
    it_ref, tr_ref = o.adam(dt=1e-3, max_iter=40)
    m = qmin.Adam(dt=1e-3, max_iter=40)
    m.set_trace(64)
    it, fm = m.minimize(f, g)
    assert it == it_ref == 40
    check(m.trace(), tr_ref, g, o)
    assert fm == pytest.approx(tr_ref[-1], rel=1e-10)


def test_adam_stops_on_tolerance():
    o, g, f = build_pair((16, 12, 10), ONE, 1, True)
    it_ref, tr_ref = o.adam(dt=2e-3, fc=0.05, max_iter=2000)
    m = qmin.Adam(dt=2e-3, fc=0.05, max_iter=2000)
    it, fm = m.minimize(f, g)
    assert it == it_ref < 2000
    assert fm <= 0.05
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-10)


@pytest.mark.parametrize("scheduled,mu,dt", [(False, 0.5, 1e-2), (True, 0.5, 1e-2), (False, 0.01, 5e-3)])
def test_nesterov_trajectory(scheduled, mu, dt):
    o, g, f = build_pair((22, 14, 12), MULTI, 0, True)
    it_ref, tr_ref = o.nesterov(dt=dt, mu=mu, scheduled=scheduled, max_iter=40)
    m = qmin.Nesterov(dt=dt, mu=mu, scheduled=scheduled, max_iter=40)
    m.set_trace(64)
    it, _ = m.minimize(f, g)
    assert it == it_ref == 40
    check(m.trace(), tr_ref, g, o)


@pytest.mark.parametrize("mem,scheduled", [(5, False), (8, False), (5, True)])
def test_lolbfgs_trajectory(mem, scheduled):
    o, g, f = build_pair((20, 16, 14), MULTI, 0, True)
    it_ref, tr_ref = o.lolbfgs(m=mem, dt=1e-2, scheduled=scheduled, max_iter=30)
    m = qmin.LoLBFGS(m=mem, dt=1e-2, scheduled=scheduled, max_iter=30)
    m.set_trace(64)
    it, _ = m.minimize(f, g)
    assert it == it_ref == 30
    check(m.trace(), tr_ref, g, o, rtol=1e-9, atol_q=1e-10)


def test_lolbfgs_all_constants_periodic():
    o, g, f = build_pair((18, 18, 18), ALL, 1, False)
    it_ref, tr_ref = o.lolbfgs(m=5, dt=1e-2, max_iter=25)
    m = qmin.LoLBFGS(m=5, dt=1e-2, max_iter=25)
    m.set_trace(32)
    m.minimize(f, g)
    check(m.trace(), tr_ref, g, o, rtol=1e-9, atol_q=1e-10)


def test_lolbfgs_continues_across_calls():
    """Two minimize calls of 12 + 18 iterations equal one of 30: the history,
    the direction and the slot counter persist on the device."""
    o, g, f = build_pair((18, 14, 12), MULTI, 0, True)
    it_ref, tr_ref = o.lolbfgs(m=5, dt=1e-2, max_iter=30)
    m = qmin.LoLBFGS(m=5, dt=1e-2, max_iter=12)
    m.set_trace(64)
    m.minimize(f, g)
    m.set_max_iterations(30)
    it, _ = m.minimize(f, g)
    assert it == 30
    check(m.trace(), tr_ref, g, o, rtol=1e-9, atol_q=1e-10)


def test_nesterov_then_fire_keeps_state_consistent():
    """Switching minimizers on one lattice: FIRE after Nesterov starts from the
    Nesterov result, as with the reference's shared model."""
    o, g, f = build_pair((16, 14, 12), MULTI, 0, True)
    o.nesterov(dt=1e-2, mu=0.5, max_iter=10)
    o.fire(max_iter=15)
    qmin.Nesterov(dt=1e-2, mu=0.5, max_iter=10).minimize(f, g)
    qmin.FIRE(max_iter=15).minimize(f, g)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("kind", ["adam", "nesterov"])
def test_fused_step_keeps_velocity_and_stored_force(kind):
    """An Adam or Nesterov iteration is one lattice pass that writes the new Q
    and the force to the other state set and leaves the FIRE velocity alone.
    FIRE, then 7 iterations (an odd number of set flips) in two calls, then FIRE
    again follows the oracle; the velocity is unchanged by the 7 iterations; and
    the stored force is the force at the Q before the last update, which is
    what the reference's force array holds after its last step
    (energyMinimizerAdam.cpp:25-54, energyMinimizerNesterovAG.cpp:75-92)."""
    o, g, f = build_pair((16, 14, 12), MULTI, 0, True)
    o.fire(max_iter=10)
    qmin.FIRE(max_iter=10).minimize(f, g)
    v_before = g.v()
    assert np.abs(v_before).max() > 0
    if kind == "adam":
        o.adam(dt=1e-3, max_iter=7)
        m = qmin.Adam(dt=1e-3, max_iter=6)
    else:
        o.nesterov(dt=1e-2, mu=0.5, max_iter=7)
        m = qmin.Nesterov(dt=1e-2, mu=0.5, max_iter=6)
    m.minimize(f, g)
    q6 = g.q()
    m.set_max_iterations(7)
    it, _ = m.minimize(f, g)
    assert it == 7
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
    assert_forces_close(g.f(), o.forces(q6))
    assert np.array_equal(g.v(), v_before)
    o.fire(max_iter=10)
    qmin.FIRE(max_iter=10).minimize(f, g)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("kind", ["adam", "nesterov"])
@pytest.mark.parametrize("precision", [qmin.QM_FP64, qmin.QM_FP32])
def test_fused_step_on_wide_tiles(kind, precision):
    """Lattices at least 128 sites wide in x and y take the 512-thread tiles
    (32 x 16 in FP64, 64 x 8 in FP32) for the Adam and Nesterov passes."""
    o, g, f = build_pair((128, 128, 6), MULTI, 0, True, precision=precision)
    if kind == "adam":
        _, tr_ref = o.adam(dt=1e-3, max_iter=12)
        m = qmin.Adam(dt=1e-3, max_iter=12)
    else:
        _, tr_ref = o.nesterov(dt=1e-2, mu=0.5, scheduled=True, max_iter=12)
        m = qmin.Nesterov(dt=1e-2, mu=0.5, scheduled=True, max_iter=12)
    m.set_trace(16)
    it, _ = m.minimize(f, g)
    assert it == 12
    if precision == qmin.QM_FP64:
        check(m.trace(), tr_ref, g, o)
    else:
        np.testing.assert_allclose(m.trace(), tr_ref, rtol=1e-4)


@pytest.mark.parametrize("kind", ["adam", "nesterov"])
def test_fused_step_without_objects(kind):
    """A periodic lattice with no objects: the boundary pass has no sites, and
    the fused pass alone finishes every site."""
    o, g, f = build_pair((18, 14, 10), ALL, 1, False)
    assert g.n_boundary() == 0
    if kind == "adam":
        _, tr_ref = o.adam(dt=1e-3, max_iter=20)
        m = qmin.Adam(dt=1e-3, max_iter=20)
    else:
        _, tr_ref = o.nesterov(dt=1e-2, mu=0.5, max_iter=20)
        m = qmin.Nesterov(dt=1e-2, mu=0.5, max_iter=20)
    m.set_trace(32)
    it, _ = m.minimize(f, g)
    assert it == 20
    check(m.trace(), tr_ref, g, o)


def test_fp32_minimizers_track_fp64():
    for name in ("adam", "nesterov", "lolbfgs"):
        o, g, f = build_pair((20, 16, 12), MULTI, 1, True, precision=qmin.QM_FP32)
        if name == "adam":
            _, tr_ref = o.adam(dt=1e-3, max_iter=20)
            m = qmin.Adam(dt=1e-3, max_iter=20)
        elif name == "nesterov":
            _, tr_ref = o.nesterov(dt=1e-2, mu=0.5, max_iter=20)
            m = qmin.Nesterov(dt=1e-2, mu=0.5, max_iter=20)
        else:
            _, tr_ref = o.lolbfgs(m=5, dt=1e-2, max_iter=20)
            m = qmin.LoLBFGS(m=5, dt=1e-2, max_iter=20)
        m.set_trace(32)
        m.minimize(f, g)
        np.testing.assert_allclose(m.trace(), tr_ref, rtol=1e-4, err_msg=name)


@pytest.mark.parametrize("kind", ["adam", "nesterov", "lolbfgs"])
def test_minimizer_moved_to_a_new_lattice_restarts(kind):
    """A minimizer used on one lattice and then on a new lattice of the same
    size (possibly at the same address once the first is destroyed) starts over
    on the new lattice: iteration 0, fresh history, as a new minimizer does."""
    make = {"adam": lambda: qmin.Adam(dt=1e-3, max_iter=10),
            "nesterov": lambda: qmin.Nesterov(dt=1e-2, mu=0.5, scheduled=True, max_iter=10),
            "lolbfgs": lambda: qmin.LoLBFGS(m=5, dt=1e-2, max_iter=10)}[kind]
    o, g, f = build_pair((16, 14, 12), MULTI, 0, True)
    m = make()
    m.minimize(f, g)
    del g
    _, g2, _ = build_pair((16, 14, 12), MULTI, 0, True, seed=99)
    _, g3, _ = build_pair((16, 14, 12), MULTI, 0, True, seed=99)
    it, fm = m.minimize(f, g2)
    it_ref, fm_ref = make().minimize(f, g3)
    assert it == it_ref == 10
    assert fm == fm_ref
    assert np.array_equal(g2.q(), g3.q())
