"""qm_fire_run_iterations_noalloc, the allocation-free FIRE entry point, through
the C ABI against qm_fire_run_iterations on the same seeded inputs: the same
kernels with the partial sums in the library's fixed scratch array instead of
the lattice's own buffer, so Q, v, F and the FIRE scalars are bit-identical.
Two calls in a row, the second on a smaller lattice (the scratch array keeps the
first call's contents)."""
import numpy as np
import pytest

import pyoracle as po
import qmin

pytestmark = pytest.mark.gpu

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)
MULTI = (2.32, 0.5, 0.3, 0.0, 0.0)


def scene(shape, precision):
    g = qmin.Lattice(*shape, precision=precision)
    g.random_init(1234, S0)
    lx, ly, lz = shape
    g.wall(2, 0, 1, 5.3, S0)
    g.wall(2, lz - 1, 1, 5.3, S0)
    g.sphere((lx / 2.0, ly / 2.0, lz / 2.0), min(shape) / 4.0, 0, 5.3, S0)
    return g


@pytest.mark.parametrize("precision", [qmin.QM_FP64, qmin.QM_FP32])
def test_noalloc_fire_equals_allocating_fire(precision):
    f = qmin.LandauDeGennesLC(A, B, C, *MULTI, semantics=1)
    for shape, n in (((48, 40, 24), 30), ((20, 18, 12), 25)):
        print("no-alloc FIRE: shape %s, %d iterations, precision %d" % (shape, n, precision), flush=True)
        ga, gb = scene(shape, precision), scene(shape, precision)
        # set-up: the object types reach the device in an ordinary (allocating) call
        fa, fb = f.compute_forces(ga), f.compute_forces(gb)
        assert np.array_equal(fa, fb)
        ma, mb = qmin.FIRE(max_iter=10 ** 6), qmin.FIRE(max_iter=10 ** 6)
        sa = ma.run_iterations(f, ga, n)
        sb = mb.run_iterations_noalloc(f, gb, n)
        assert sa.iterations == sb.iterations == n
        for name in ("forceMax", "power", "alpha", "deltaT", "scaling", "nPositive"):
            assert getattr(sa, name) == getattr(sb, name), name
        assert np.array_equal(ga.q(), gb.q())
        assert np.array_equal(ga.v(), gb.v())
        assert np.array_equal(ga.f(), gb.f())
        # a second call continues the same run
        sa = ma.run_iterations(f, ga, 7)
        sb = mb.run_iterations_noalloc(f, gb, 7)
        assert sa.forceMax == sb.forceMax and sa.iterations == sb.iterations == n + 7
        assert np.array_equal(ga.q(), gb.q())


def test_noalloc_fire_refuses_a_lattice_beyond_the_scratch_array():
    """1024 x 256 x 260 (FP32, 68M sites) in 32 x 8 tiles and 4-plane z chunks
    is 1024 x 65 = 66,560 work-groups per pass, more than the 65,536 slots: the
    call fails before any launch (the wall uploads the site types)."""
    f = qmin.LandauDeGennesLC(A, B, C, *MULTI, semantics=1)
    g = qmin.Lattice(1024, 256, 260, precision=qmin.QM_FP32)
    g.set_tile(0)
    g.set_zchunk(4)
    g.wall(2, 0, 1, 5.3, S0)
    m = qmin.FIRE(max_iter=100)
    with pytest.raises(qmin.QminError, match="partial-sum slots"):
        m.run_iterations_noalloc(f, g, 3)
    assert m.state().iterations == 0
