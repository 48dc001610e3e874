"""Ensembles of uneven lattices minimized together (batched launches replayed
from a hipGraph) against the oracle, lattice by lattice."""
import numpy as np
import pytest

import pyoracle as po
import qmin

pytestmark = pytest.mark.gpu

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)

# (shape, seed, sphere (centre, radius, type, W) or None)
SPECS = [
    ((24, 32, 16), 11, ((12.0, 16.0, 8.0), 5.0, 0, 5.3)),
    ((32, 24, 24), 12, ((10.0, 12.0, 11.5), 6.5, 1, 0.8)),
    ((48, 24, 20), 13, None),
    ((24, 24, 48), 14, ((12.0, 12.0, 30.0), 4.0, 0, 12.0)),
    ((17, 9, 11), 15, ((8.0, 4.0, 5.0), 2.5, 0, 2.0)),
    ((64, 32, 24), 16, ((40.0, 10.0, 12.0), 7.0, 1, 20.0)),
]


def make(spec, lib_side, precision=qmin.QM_FP64):
    shape, seed, sph = spec
    if lib_side:
        x = qmin.Lattice(*shape, precision=precision)
    else:
        x = po.OracleSystem(*shape)
    x.random_init(seed, S0)
    if sph is not None:
        c, r, t, w = sph
        x.sphere(c, r, t, w, S0)
    return x


@pytest.mark.parametrize("graph_iterations", [8, 32])
def test_ensemble_fire_matches_oracle_per_lattice(graph_iterations):
    lats = [make(s, True) for s in SPECS]
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.0, 0.0, semantics=1)
    fire = qmin.FIRE(fc=1e-6, max_iter=150)
    ens = qmin.Ensemble(lats)
    its, fms = ens.fire_minimize(fire, force, graph_iterations)
    for spec, lat, it, fm in zip(SPECS, lats, its, fms):
        o = make(spec, False)
        o.set_params(A, B, C, 2.32, semantics=1)
        it_ref, tr = o.fire(fc=1e-6, max_iter=150)
        assert it == it_ref, spec
        assert fm == pytest.approx(tr[-1, 0], rel=1e-9), spec
        np.testing.assert_allclose(lat.q(), o.q(), rtol=0, atol=1e-10, err_msg=str(spec))


def test_ensemble_lattices_stop_on_their_own_tolerance():
    """Different lattices converge after different numbers of iterations; each
    stops exactly where its own minimize() would."""
    lats = [make(s, True) for s in SPECS[:4]]
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=0)
    fire = qmin.FIRE(fc=2e-3, max_iter=3000)
    its, fms = qmin.Ensemble(lats).fire_minimize(fire, force, 16)
    assert len(set(its.tolist())) > 1
    for spec, it, fm in zip(SPECS[:4], its, fms):
        o = make(spec, False)
        o.set_params(A, B, C, 2.32, 0.5, 0.3, semantics=0)
        it_ref, tr = o.fire(fc=2e-3, max_iter=3000)
        assert it == it_ref and fm <= 2e-3
        assert fm == pytest.approx(tr[-1, 0], rel=1e-9)


def test_ensemble_fp32_tracks_single_lattice_fp32():
    lats = [make(s, True, qmin.QM_FP32) for s in SPECS]
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.0, 0.0, semantics=1)
    its, fms = qmin.Ensemble(lats).fire_minimize(qmin.FIRE(fc=1e-7, max_iter=120), force, 16)
    for spec, lat, fm in zip(SPECS, lats, fms):
        single = make(spec, True, qmin.QM_FP32)
        st = qmin.FIRE(fc=1e-7, max_iter=120).minimize(force, single)
        assert fm == pytest.approx(st.forceMax, rel=1e-4)
        np.testing.assert_allclose(lat.q(), single.q(), atol=1e-5)


def test_ensemble_autotune_leaves_states_and_results_unchanged():
    """The ensemble tile autotuner leaves every lattice as it was, and FIRE on
    the tuned shape stops where each lattice's own minimize() does (per-site
    forces do not depend on the tile; the reductions only by rounding)."""
    lats = [make(s, True) for s in SPECS]
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=1)
    q0 = [l.q() for l in lats]
    ens = qmin.Ensemble(lats)
    tile, zchunk, ms = ens.autotune(force)
    assert tile in (qmin.TILE_PER_LATTICE, 0, 1, 2, 3) and zchunk in (8, 16, 32, 64) and ms > 0
    for l, q in zip(lats, q0):
        # equal_nan: the object site at an integer sphere centre has no direction
        # (0/0 in the reference's Q^B as well); it never enters a force
        assert np.array_equal(l.q(), q, equal_nan=True)
    its, fms = ens.fire_minimize(qmin.FIRE(fc=1e-6, max_iter=150), force, 16)
    for spec, lat, it, fm in zip(SPECS, lats, its, fms):
        o = make(spec, False)
        o.set_params(A, B, C, 2.32, 0.5, 0.3, semantics=1)
        it_ref, tr = o.fire(fc=1e-6, max_iter=150)
        assert it == it_ref, spec
        assert fm == pytest.approx(tr[-1, 0], rel=1e-9), spec
        np.testing.assert_allclose(lat.q(), o.q(), rtol=0, atol=1e-10, err_msg=str(spec))


def test_ensemble_with_a_tile_per_lattice_matches_oracle():
    """QM_TILE_PER_LATTICE: lattices 17, 24 and 32 sites wide run 32 x 8 tiles,
    48 and 64 wide 64 x 8 (two launches of the batched pass per iteration); each
    lattice still stops where its own oracle run does."""
    lats = [make(s, True) for s in SPECS]
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=1)
    ens = qmin.Ensemble(lats)
    ens.set_tile(qmin.TILE_PER_LATTICE, 16)
    its, fms = ens.fire_minimize(qmin.FIRE(fc=1e-6, max_iter=150), force, 16)
    for spec, lat, it, fm in zip(SPECS, lats, its, fms):
        o = make(spec, False)
        o.set_params(A, B, C, 2.32, 0.5, 0.3, semantics=1)
        it_ref, tr = o.fire(fc=1e-6, max_iter=150)
        assert it == it_ref, spec
        assert fm == pytest.approx(tr[-1, 0], rel=1e-9), spec
        np.testing.assert_allclose(lat.q(), o.q(), rtol=0, atol=1e-10, err_msg=str(spec))


def test_ensemble_set_tile_rejects_bad_values():
    lats = [make(s, True) for s in SPECS[:2]]
    ens = qmin.Ensemble(lats)
    for tile, zc in ((4, 0), (-3, 0), (0, 2), (0, -1)):
        with pytest.raises(qmin.QminError):
            ens.set_tile(tile, zc)
