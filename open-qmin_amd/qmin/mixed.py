"""Mixed-precision FIRE: the lattice's state and stencils in FP32 (with FP64
reductions) while FP32 still makes progress, then an on-device conversion of
the state to FP64 and a polish by the same minimizer (dt, alpha and N+ kept) to
the requested tolerance.

FP32 rounding puts a floor under forceMax (the reference's |F|_2 / N), which
falls with the lattice size: about 2.2e-10 at 256^3 with walls and a sphere.
FP32 runs in chunks of FIRE iterations and hands over to FP64 once forceMax
falls below switch_fc or a chunk's smallest forceMax is no longer a factor
`stall` below the previous chunk's.
"""
import numpy as np

from . import FIRE, QM_FP32, QM_FP64


def fire_mixed(force, lattice, fc, switch_fc=0.0, chunk=2000, stall=0.9, max_iter=10 ** 6, **fire_kw):
    """FIRE to forceMax < fc in mixed precision.  fire_kw are FIRE's parameters
    (dt, dt_max, ...).  The lattice is converted to FP32 first if it is not.
    Returns (the FIRE minimizer, its state, the FP32 iterations)."""
    if lattice.precision != QM_FP32:
        lattice.set_precision(QM_FP32)
    m = FIRE(fc=max(fc, switch_fc), max_iter=min(chunk, max_iter), **fire_kw)
    m.set_trace(chunk)
    best_prev = np.inf
    while True:
        st = m.minimize(force, lattice)
        if st.converged or st.nanDetected or st.iterations >= max_iter:
            break
        best = float(m.trace(st.iterations)[-chunk:, 0].min())
        if best > stall * best_prev:
            break
        best_prev = best
        m.set_max_iterations(min(st.iterations + chunk, max_iter))
    fp32_iterations = st.iterations
    lattice.set_precision(QM_FP64)
    m.set_force_cutoff(fc)
    m.set_max_iterations(max_iter)
    st = m.minimize(force, lattice)
    return m, st, fp32_iterations
