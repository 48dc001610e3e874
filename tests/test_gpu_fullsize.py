"""The HIP path at the benchmark's full size (256^3), through properties that do
not depend on size (the oracle checks the same kernels at small sizes):

  * translation equivariance: forces of a periodically shifted state are the
    shifted forces, bit for bit (the per-site arithmetic does not depend on where
    a site falls in a tile, a z chunk or the grid);
  * a uniform nematic at S0 is stationary (no force, FIRE stops at once);
  * with the phase term off, an L1 plane wave along any axis has the closed-form
    force -4 L1 (1 - cos k) Q (one-constant and composed multi-constant paths),
    also at 512^3, where byte offsets into a field pass 2^32;
  * the energy is extensive: a 2x2x2 periodic copy of a 128^3 state has 8 times
    its energy;
  * FIRE keeps a 2x2x2 periodic copy exactly periodic, and follows the 128^3
    run's forceMax to reduction rounding;
  * at 768^3 (453M sites, 5n = 2.26e9 values per field, so element offsets
    pass 2^31) FIRE and then Nesterov keep a state that varies only along z
    uniform in every z plane, and follow the same run on an 8 x 8 x 768 lattice.
"""
import numpy as np
import pytest

import pyoracle as po
import qmin

pytestmark = pytest.mark.gpu

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)
L = 256


def grid(q, shape):
    lx, ly, lz = shape
    return q.reshape(lz, ly, lx, 5)


@pytest.mark.parametrize("precision,params", [(qmin.QM_FP64, (2.32, 0.5, 0.3)), (qmin.QM_FP32, (2.32, 0.0, 0.0))])
def test_translation_equivariance_256(precision, params):
    lat = qmin.Lattice(L, L, L, precision=precision)
    lat.random_init(1234, S0)
    f = qmin.LandauDeGennesLC(A, B, C, *params, semantics=qmin.SEMANTICS_REF_GPU)
    F0 = grid(f.compute_forces(lat), (L, L, L))
    q = grid(lat.q(), (L, L, L))
    shift = (7, 5, 3)   # (z, y, x) sites
    lat.set_field(qmin.FIELD_Q, np.roll(q, shift, axis=(0, 1, 2)).reshape(-1, 5))
    F1 = grid(f.compute_forces(lat), (L, L, L))
    assert np.array_equal(F1, np.roll(F0, shift, axis=(0, 1, 2)))


def test_uniform_nematic_is_stationary_256():
    lat = qmin.Lattice(L, L, L)
    lat.uniform_init((0.3, -0.5, 0.8), S0)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    F = f.compute_forces(lat)
    assert np.abs(F).max() < 1e-13
    st = qmin.FIRE(fc=1e-12, max_iter=100).minimize(f, lat)
    assert st.iterations == 1 and st.forceMax < 1e-12


@pytest.mark.parametrize("multi", [False, True])
def test_l1_plane_wave_closed_form_256(multi):
    """Q(r) = eps sin(k r_axis) e, k = 2 pi 5 / 256, phase off: 6Q - (sum of
    neighbours) = 2 (1 - cos k) Q, and the L1 shaping followed by the GPU path's
    metric correction is twice the identity, so F = -4 L1 (1 - cos k) Q."""
    L1 = 2.32
    f = qmin.LandauDeGennesLC(0.0, 0.0, 0.0, L1, multi=multi, semantics=qmin.SEMANTICS_REF_GPU)
    lat = qmin.Lattice(L, L, L)
    e = np.array([0.3, -0.2, 0.5, 0.1, -0.4])
    k = 2.0 * np.pi * 5 / L
    wave = np.sin(k * np.arange(L))
    for axis in range(3):
        shape = [1, 1, 1]
        shape[2 - axis] = L                     # (z, y, x) grid axes
        q = np.broadcast_to((0.01 * wave).reshape(shape)[..., None] * e, (L, L, L, 5)).reshape(-1, 5)
        lat.set_field(qmin.FIELD_Q, np.ascontiguousarray(q))
        F = f.compute_forces(lat)
        want = -4.0 * L1 * (1.0 - np.cos(k)) * q
        assert np.abs(F - want).max() <= 1e-13 * L1 * np.abs(q).max(), axis


def test_l1_plane_wave_closed_form_512():
    """The same closed form on a 512^3 lattice (134M sites): each FP64 field then
    spans 5.4 GB, so byte offsets pass 2^32 (element offsets reach 15n = 2.0e9,
    just below 2^31; the 768^3 test below passes that too)."""
    n = 512
    L1 = 2.32
    f = qmin.LandauDeGennesLC(0.0, 0.0, 0.0, L1, multi=True, semantics=qmin.SEMANTICS_REF_GPU)
    lat = qmin.Lattice(n, n, n)
    e = np.array([0.3, -0.2, 0.5, 0.1, -0.4])
    k = 2.0 * np.pi * 7 / n
    wave = 0.01 * np.sin(k * np.arange(n))
    q = np.empty((n, n, n, 5))
    q[...] = wave[:, None, None, None] * e          # varies along z, the slowest axis
    lat.set_field(qmin.FIELD_Q, q.reshape(-1, 5))
    F = f.compute_forces(lat).reshape(n, n, n, 5)
    c = -4.0 * L1 * (1.0 - np.cos(k))
    # rounding of 6Q - (sum of neighbours) scales with the input, not with the
    # (1 - cos k)-small result; a wrong site would be off by ~|c q|, ~1e11 times this bound
    scale = L1 * np.abs(q).max()
    for z in range(0, n, 64):                      # plane by plane, to bound host temporaries
        assert np.abs(F[z:z + 64] - c * q[z:z + 64]).max() <= 1e-13 * scale, z


def test_energy_is_extensive_256():
    small = qmin.Lattice(128, 128, 128)
    small.random_init(99, S0)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    e_small, c_small = f.compute_energy(small)
    big = qmin.Lattice(L, L, L)
    big.set_field(qmin.FIELD_Q, np.tile(grid(small.q(), (128, 128, 128)), (2, 2, 2, 1)).reshape(-1, 5))
    e_big, c_big = f.compute_energy(big)
    np.testing.assert_allclose(e_big, 8 * e_small, rtol=1e-12)
    np.testing.assert_allclose(c_big, 8 * c_small, rtol=1e-11, atol=1e-9 * abs(e_big))


def test_fire_keeps_a_periodic_copy_periodic_256():
    small = qmin.Lattice(128, 128, 128)
    small.random_init(7, S0)
    q_small = grid(small.q(), (128, 128, 128))
    big = qmin.Lattice(L, L, L)
    big.set_field(qmin.FIELD_Q, np.tile(q_small, (2, 2, 2, 1)).reshape(-1, 5))
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    ms, mb = qmin.FIRE(max_iter=30), qmin.FIRE(max_iter=30)
    ms.set_trace(32)
    mb.set_trace(32)
    ms.minimize(f, small)
    mb.minimize(f, big)
    qb = grid(big.q(), (L, L, L))
    h = 128
    assert np.array_equal(qb[:h, :h, :h], qb[h:, :h, :h])
    assert np.array_equal(qb[:h, :h, :h], qb[:h, h:, h:])
    assert np.array_equal(qb[:h, :h, :h], qb[h:, h:, h:])
    # forceMax = |F| / N: the copy has sqrt(8) |F| over 8 N
    np.testing.assert_allclose(mb.trace()[:, 0] * np.sqrt(8.0), ms.trace()[:, 0], rtol=1e-10)
    np.testing.assert_allclose(qb[:h, :h, :h], grid(small.q(), (128, 128, 128)), rtol=0, atol=1e-11)


@pytest.mark.timeout(300)
def test_fire_then_nesterov_on_a_z_column_768():
    """A twisted nematic whose director and order vary only along z, on a 768^3
    FP64 lattice: each field holds 5n = 2.26e9 values, so the element offsets of
    its upper planes pass 2^31 in Q, F, the FIRE velocity and Nesterov's
    alternate sequence.  Ten FIRE iterations and then five Nesterov iterations
    must leave every z plane uniform, bit for bit (the same per-site arithmetic
    on the same inputs), and follow the same runs on an 8 x 8 x 768 lattice:
    the big lattice holds 9216 copies of the small one's sites, so 9216 times its
    energy and 1/96 of its forceMax = |F| / N.  A site skipped or read from a
    wrapped 32-bit offset breaks the plane uniformity outright."""
    n = 768
    copies = n * n // 64
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    z = np.arange(n)
    th = 2.0 * np.pi * 2 * z / n
    col = np.stack([np.cos(th), np.sin(th), np.zeros(n), S0 * (1.0 + 0.2 * np.sin(2.0 * np.pi * 3 * z / n))], axis=1)
    small = qmin.Lattice(8, 8, n)
    small.director_init(np.repeat(col, 64, axis=0))     # site = x + 8 y + 64 z
    q0 = small.q().reshape(n, 64, 5)
    big = qmin.Lattice(n, n, n)
    q = np.empty((n, n * n, 5))
    q[...] = q0[:, :1, :]
    big.set_field(qmin.FIELD_Q, q.reshape(-1, 5))
    del q

    def run(m_small, m_big, cap):
        m_small.set_trace(cap)
        m_big.set_trace(cap)
        m_small.minimize(f, small)
        m_big.minimize(f, big)
        # forceMax per iteration: column 0 of FIRE's trace rows, the trace itself otherwise
        ts, tb = (t[:, 0] if t.ndim == 2 else t for t in (m_small.trace(), m_big.trace()))
        assert len(ts) == len(tb) > 0
        np.testing.assert_allclose(tb * np.sqrt(copies), ts, rtol=1e-10)
        e_small, _ = f.compute_energy(small)
        e_big, _ = f.compute_energy(big)
        np.testing.assert_allclose(e_big, copies * e_small, rtol=1e-11)

    run(qmin.FIRE(max_iter=10), qmin.FIRE(max_iter=10), 16)
    run(qmin.Nesterov(dt=1e-3, mu=0.01, max_iter=5), qmin.Nesterov(dt=1e-3, mu=0.01, max_iter=5), 8)

    qs = small.q().reshape(n, 64, 5)
    assert np.array_equal(qs, np.broadcast_to(qs[:, :1], qs.shape))
    assert np.abs(qs - q0).max() > 1e-6                  # the runs moved the state
    qb = big.q().reshape(n, n * n, 5)
    for z0 in range(0, n, 32):                          # plane slabs bound host temporaries
        slab = qb[z0:z0 + 32]
        assert np.array_equal(slab, np.broadcast_to(slab[:, :1], slab.shape)), z0
        assert np.abs(slab[:, 0] - qs[z0:z0 + 32, 0]).max() <= 1e-11, z0
