"""The workload configurations on the MI355X against the CPU oracle.

* Full elasticity (L1-L4 and L6, a uniform E field, four homeotropic spheres;
  the 384^3 workload scaled to 24^3 and 40^3 so that the oracle finishes in
  seconds): FIRE, Adam, Nesterov and online L-BFGS (memory 8), each in FP64 and
  in the mixed form (an FP32 run, then the device conversion and an FP64 run
  from its state).  FP64 traces within 1e-10 relative per iteration (L-BFGS
  1e-9), Q within 1e-11; the FP32 phase within 1e-4 of the FP64 trace; the FP64
  phase after the conversion, started on the oracle from the device's state,
  within the FP64 bounds.  FIRE to forceMax < 1e-8 in both forms: energy within
  1e-9 relative of the oracle's (FP64 stops at the oracle's iteration).
* The 256^3 workload with walls and a sphere at its full size: forces under both
  metric semantics within 1e-12 of the force scale, three FIRE iterations
  (trace within 1e-10, Q within 1e-11), with the oracle on all host cores.
"""
import os

import numpy as np
import pytest

import pyoracle as po
import qmin
from qmin import workloads
from test_gpu_parity import assert_forces_close

pytestmark = pytest.mark.gpu

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)
ALL = (2.32, 1.0, 0.5, 0.2, 0.3)
THREADS = min(16, os.cpu_count() or 1)


def full_elasticity_pair(L, precision=qmin.QM_FP64):
    """workloads.full_elasticity on the device and the same scene on the oracle."""
    g, f = workloads.full_elasticity(qmin, L, precision)
    o = po.OracleSystem(L, L, L)
    o.threads(THREADS)
    o.random_init(1234, S0)
    rng = np.random.default_rng(384)
    for _ in range(4):
        c = tuple(float(v) for v in rng.uniform(0, L, 3))
        o.sphere(c, float(rng.integers(12, 25)) * L / 384, 0, float(rng.uniform(1, 10)), S0)
    o.set_params(A, B, C, *ALL, multi=True, semantics=1)
    o.set_efield((0.0, 0.0, 0.5), 1.0, 1.0, 1.0)
    q = o.q() if precision == qmin.QM_FP64 else o.q().astype(np.float32).astype(np.float64)
    assert np.array_equal(g.types(), o.types()) and np.array_equal(g.q(), q)
    return o, g, f


MINIMIZERS = {
    "fire": (lambda o, n: o.fire(dt=5e-4, dt_max=2e-2, max_iter=n),
             lambda n: qmin.FIRE(dt=5e-4, dt_max=2e-2, max_iter=n), 1e-10),
    "adam": (lambda o, n: o.adam(dt=1e-3, max_iter=n), lambda n: qmin.Adam(dt=1e-3, max_iter=n), 1e-10),
    "nesterov": (lambda o, n: o.nesterov(dt=1e-2, mu=0.01, max_iter=n),
                 lambda n: qmin.Nesterov(dt=1e-2, mu=0.01, max_iter=n), 1e-10),
    "lolbfgs": (lambda o, n: o.lolbfgs(m=8, dt=1e-4, tau=10.0, scheduled=True, max_iter=n),
                lambda n: qmin.LoLBFGS(m=8, dt=1e-4, tau=10.0, scheduled=True, max_iter=n), 1e-9),
}


def _trace(m):
    t = m.trace()
    return t[:, 0] if t.ndim == 2 else t


@pytest.mark.parametrize("name", list(MINIMIZERS))
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_full_elasticity_minimizers(name, mode):
    print("full elasticity 40^3: %s %s" % (name, mode), flush=True)
    run_oracle, make, rtol = MINIMIZERS[name]
    prec = qmin.QM_FP64 if mode == "fp64" else qmin.QM_FP32
    o, g, f = full_elasticity_pair(40, prec)
    n = 30
    it_ref, tr_ref = run_oracle(o, n)
    tr_ref = tr_ref[:, 0] if tr_ref.ndim == 2 else tr_ref
    m = make(n)
    m.set_trace(n + 8)
    m.minimize(f, g)
    tr = _trace(m)
    assert len(tr) == it_ref == n
    if mode == "fp64":
        np.testing.assert_allclose(tr, tr_ref, rtol=rtol)
        np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)
        return
    np.testing.assert_allclose(tr, tr_ref, rtol=1e-4)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-5)
    # the FP64 phase from the converted state: the oracle starts from the same Q and v
    g.set_precision(qmin.QM_FP64)
    o.set_q(g.q())
    o.set_v(g.v())
    it_ref, tr_ref = run_oracle(o, 10)
    tr_ref = tr_ref[:, 0] if tr_ref.ndim == 2 else tr_ref
    m = make(10)
    m.set_trace(16)
    m.minimize(f, g)
    np.testing.assert_allclose(_trace(m), tr_ref, rtol=rtol)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_full_elasticity_fire_to_tolerance(mode):
    """FIRE to forceMax < 1e-8 on the 24^3 full-elasticity scene: the converged
    energy within 1e-9 relative of the oracle's.  Mixed: FP32 until 2e-8 (near
    its rounding floor on 13,824 sites), then the FP64 polish to 1e-8."""
    print("full elasticity 24^3 FIRE to 1e-8: %s" % mode, flush=True)
    prec = qmin.QM_FP64 if mode == "fp64" else qmin.QM_FP32
    o, g, f = full_elasticity_pair(24, prec)
    it_ref, tr_ref = o.fire(dt=5e-4, dt_max=2e-2, fc=1e-8, max_iter=20000)
    assert tr_ref[-1, 0] < 1e-8
    m = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=1e-8 if mode == "fp64" else 2e-8, max_iter=20000)
    st = m.minimize(f, g)
    if mode == "mixed":
        g.set_precision(qmin.QM_FP64)
        m.set_force_cutoff(1e-8)
        st = m.minimize(f, g)
    assert st.converged and st.forceMax < 1e-8
    if mode == "fp64":
        assert st.iterations == it_ref
    e_ref, _ = o.energy()
    e_got, _ = f.compute_energy(g)
    assert abs(e_got - e_ref) <= 1e-9 * abs(e_ref), (e_got, e_ref)


def test_walls_sphere_256_full_size():
    """The 256^3 scene (planar walls at z = 0 and 255, a homeotropic sphere of
    radius 20 at the centre, L1-L3) at full size: forces under both metric
    semantics and three FIRE iterations against the oracle."""
    L = 256
    print("256^3 walls + sphere against the oracle (%d threads)" % THREADS, flush=True)
    o = po.OracleSystem(L, L, L)
    o.threads(THREADS)
    g = qmin.Lattice(L, L, L)
    for s in (o, g):
        s.random_init(1234, S0)
        s.wall(2, 0, 1, 5.3, S0)
        s.wall(2, L - 1, 1, 5.3, S0)
        s.sphere((128.0, 128.0, 128.0), 20.0, 0, 5.3, S0)
    assert np.array_equal(g.types(), o.types())
    for sem in (0, 1):
        o.set_params(A, B, C, 2.32, 0.5, 0.3, multi=True, semantics=sem)
        f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=sem)
        ref = o.forces()
        got = f.compute_forces(g)
        assert_forces_close(got, ref)
        del ref, got
    it_ref, tr_ref = o.fire(max_iter=3)
    m = qmin.FIRE(max_iter=3)
    m.set_trace(4)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


def test_fire_mixed_hands_over_when_fp32_stalls():
    """qmin.mixed.fire_mixed with no switch tolerance: FP32 runs until its
    forceMax stops falling (its rounding floor), then FP64 polishes to 1e-11.
    On the 48^3 walls-and-sphere scene the FP32 phase ends above 1e-11 and the
    converged energy is within 1e-9 relative of an all-FP64 run's."""
    from qmin import mixed
    from bench import build_workload, fire_params
    print("mixed FIRE with a stall hand-over, 48^3", flush=True)
    g64, f = build_workload(qmin, 48, qmin.QM_FP64)
    st64 = qmin.FIRE(fc=1e-11, max_iter=400000, **fire_params()).minimize(f, g64)
    assert st64.converged
    e64, _ = f.compute_energy(g64)
    gmx, _ = build_workload(qmin, 48, qmin.QM_FP32)
    m, st, fp32_its = mixed.fire_mixed(f, gmx, 1e-11, chunk=1000, max_iter=400000, **fire_params())
    assert st.converged and st.forceMax < 1e-11
    assert 0 < fp32_its < st.iterations
    assert gmx.precision == qmin.QM_FP64
    emx, _ = f.compute_energy(gmx)
    assert abs(emx - e64) <= 1e-9 * abs(e64), (emx, e64)
