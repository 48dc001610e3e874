"""The reference's own input files and the error paths of the C ABI.

tests/golden/assets/ holds three input files that ship with the reference: two
boundary files (assets/oneSphereExample.txt: one homeotropic sphere, W 5.3,
S0 0.53, 4224 sites; assets/twoSphereExample.txt: a homeotropic and a
degenerate-planar sphere of 540 sites each), in the format multirankSimulation::
createBoundaryFromFile reads (multirankSimulationBoundaries.cpp:41-96): a line
with the number of objects, then per object a line "type W S0 count" and count
lines "x y z Qxx Qxy Qxz Qyy Qyz"; and my_init_file_x0y1z0.txt (gzipped), an
initial configuration in the state-file format (one rank's share of a 50^3
cholesteric helix).

CPU: the files read by read_boundary_file below build the same site types
in the oracle as in the reference's compiled cubicLattice.  GPU: the library's
own reader (qm_lattice_boundary_file) against the oracle, bit for bit, and FIRE
from there; the state file (saveState / loadState, multirankSimulation.cpp:
573-665) round trip; and the error codes a caller gets for a missing file, a
site outside the lattice, and a NaN state, also for one lattice of an ensemble.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import REPO
from test_gpu_parity import assert_forces_close

ASSETS = os.path.join(REPO, "tests", "golden", "assets")
A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)
FILES = {"oneSphereExample.txt": (50, 50, 50), "twoSphereExample.txt": (40, 40, 40)}


def read_boundary_file(path):
    """[(btype, W, S0, xyz[n][3], qb[n][5])] in file order."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    count = int(lines[0].split()[0])
    out, i = [], 1
    for _ in range(count):
        bt, W, s0, n = lines[i].split()[:4]
        n = int(n)
        rows = np.array([[float(v) for v in l.split()[:8]] for l in lines[i + 1:i + 1 + n]])
        out.append((int(bt), float(W), float(s0), rows[:, :3].astype(np.int32), rows[:, 3:8].copy()))
        i += 1 + n
    return out


def oracle_from_file(name, seed=1234):
    shape = FILES[name]
    o = po.OracleSystem(*shape)
    o.random_init(seed, S0)
    for bt, W, s0, xyz, qb in read_boundary_file(os.path.join(ASSETS, name)):
        o.add_object(xyz, qb, bt, W, s0)
    return o


@pytest.mark.skipif(not po.ref_available(), reason="needs oracle/_ref (reference checkout)")
@pytest.mark.parametrize("name", sorted(FILES))
def test_boundary_file_types_match_compiled_reference(name):
    """The objects of the reference's example files give the same site types in
    the oracle as in the compiled cubicLattice::createBoundaryObject."""
    o = oracle_from_file(name)
    shape = FILES[name]
    r = po.RefModel(*shape)
    objs = read_boundary_file(os.path.join(ASSETS, name))
    for bt, W, s0, xyz, qb in objs:
        sites = (xyz[:, 0] + shape[0] * (xyz[:, 1] + shape[1] * xyz[:, 2])).astype(np.int32)
        r.add_object(sites, bt, W, s0)
    assert np.array_equal(r.types(), o.types())
    assert len(objs) == {"oneSphereExample.txt": 1, "twoSphereExample.txt": 2}[name]
    q = o.q()
    for k, (bt, W, s0, xyz, qb) in enumerate(objs):
        sites = xyz[:, 0] + shape[0] * (xyz[:, 1] + shape[1] * xyz[:, 2])
        assert np.all(o.types()[sites] == k + 1)
        assert np.array_equal(q[sites], qb)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FILES))
def test_boundary_file_on_the_device_matches_oracle(name):
    import qmin
    print("boundary file %s" % name, flush=True)
    o = oracle_from_file(name)
    g = qmin.Lattice(*FILES[name])
    g.random_init(1234, S0)
    g.boundary_file(os.path.join(ASSETS, name))
    assert np.array_equal(g.types(), o.types())
    assert np.array_equal(g.q(), o.q())
    o.set_params(A, B, C, 2.32, 0.5, 0.3, multi=True, semantics=1)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=1)
    it_ref, tr_ref = o.fire(max_iter=20)
    m = qmin.FIRE(max_iter=20)
    m.set_trace(32)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_state_file_round_trip_and_reference_format(tmp_path):
    """saveState writes "x y z Qxx Qxy Qxz Qyy Qyz type Smax" with six decimals;
    loadState reads Q back (types come from the objects, as in the reference,
    whose loader leaves them alone, multirankSimulation.cpp:606); a file in that
    format written here loads exactly, and FIRE from the loaded state follows the
    oracle started from the same Q."""
    import qmin
    o = oracle_from_file("twoSphereExample.txt", seed=5)
    g = qmin.Lattice(40, 40, 40)
    g.random_init(5, S0)
    g.boundary_file(os.path.join(ASSETS, "twoSphereExample.txt"))
    path = str(tmp_path / "state_x0y0z0.txt")
    g.save_state(path)
    rows = np.loadtxt(path)
    assert rows.shape == (40 ** 3, 10)
    g2 = qmin.Lattice(40, 40, 40)
    g2.boundary_file(os.path.join(ASSETS, "twoSphereExample.txt"))
    g2.load_state(path)
    site = (rows[:, 0] + 40 * (rows[:, 1] + 40 * rows[:, 2])).astype(int)
    want = np.empty((site.size, 5))
    want[site] = rows[:, 3:8]
    assert np.array_equal(g2.q(), want)
    np.testing.assert_allclose(g2.q(), g.q(), atol=5e-7)
    assert np.array_equal(g2.types(), g.types())
    # a file written in the reference's format, full precision
    q = o.q()
    with open(str(tmp_path / "ref_x0y0z0.txt"), "w") as fh:
        for i in range(o.n):
            x, y, z = i % 40, (i // 40) % 40, i // 1600
            fh.write("%d %d %d %.17g %.17g %.17g %.17g %.17g %d %.6f\n" % (x, y, z, *q[i], o.types()[i], 0.5))
    g3 = qmin.Lattice(40, 40, 40)
    g3.boundary_file(os.path.join(ASSETS, "twoSphereExample.txt"))
    g3.load_state(str(tmp_path / "ref_x0y0z0.txt"))
    assert np.array_equal(g3.q(), q)
    o.set_params(A, B, C, 2.32, semantics=1)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=1)
    _, tr_ref = o.fire(max_iter=15)
    m = qmin.FIRE(max_iter=15)
    m.set_trace(16)
    m.minimize(f, g3)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)


def helix_q(z, lz=50, S=0.53):
    """Q = (3S/2)(n n - I/3) of the cholesteric helix n = (cos 2 pi z/lz,
    sin 2 pi z/lz, 0) that the reference's tools/initHelper.py wrote into
    my_init_file_*.txt (pitch lz along z, S = 0.53)."""
    n = np.stack([np.cos(2 * np.pi * z / lz), np.sin(2 * np.pi * z / lz), 0 * z], 1)
    Q = 1.5 * S * (n[:, :, None] * n[:, None, :] - np.eye(3) / 3)
    return np.stack([Q[:, 0, 0], Q[:, 0, 1], Q[:, 0, 2], Q[:, 1, 1], Q[:, 1, 2]], 1)


def init_file_rows():
    import gzip
    with gzip.open(os.path.join(ASSETS, "my_init_file_x0y1z0.txt.gz"), "rt") as fh:
        return np.loadtxt(fh)


def test_init_file_fixture_is_the_helix():
    """my_init_file_x0y1z0.txt (gzipped here) is rank (0,1,0) of a 2x2x1 split of
    a 50^3 box: 25 x 25 x 50 sites in global coordinates (x 0..24, y 25..49),
    type 0, the helix's Q to six decimals."""
    rows = init_file_rows()
    assert rows.shape == (31250, 10)
    assert rows[:, 0].min() == 0 and rows[:, 0].max() == 24 and rows[:, 1].min() == 25 and rows[:, 1].max() == 49
    assert np.all(rows[:, 8] == 0)
    assert np.abs(rows[:, 3:8] - helix_q(rows[:, 2])).max() <= 5e-7


@pytest.mark.gpu
def test_load_state_of_the_reference_init_file(tmp_path):
    """qm_lattice_load_state on the reference's initial-configuration file: the
    sites it lists take its values (global coordinates, as loadState reads them
    on one rank), every other site keeps its Q; FIRE from there follows the
    oracle started from the same Q."""
    import qmin
    import gzip
    rows = init_file_rows()
    path = str(tmp_path / "my_init_file_x0y1z0.txt")
    with gzip.open(os.path.join(ASSETS, "my_init_file_x0y1z0.txt.gz"), "rb") as src, open(path, "wb") as dst:
        dst.write(src.read())
    g = qmin.Lattice(50, 50, 50)
    g.random_init(77, S0)
    before = g.q()
    g.load_state(path)
    q = g.q()
    site = (rows[:, 0] + 50 * (rows[:, 1] + 50 * rows[:, 2])).astype(int)
    assert np.array_equal(q[site], rows[:, 3:8])
    rest = np.ones(g.n, bool)
    rest[site] = False
    assert np.array_equal(q[rest], before[rest])
    assert np.abs(q[site] - helix_q(rows[:, 2])).max() <= 5e-7
    o = po.OracleSystem(50, 50, 50)
    o.set_q(q)
    o.set_params(A, B, C, 2.32, semantics=1)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=1)
    _, tr_ref = o.fire(max_iter=12)
    m = qmin.FIRE(max_iter=12)
    m.set_trace(16)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_error_codes_for_bad_files(tmp_path):
    import qmin
    g = qmin.Lattice(12, 12, 12)
    lib = qmin.library()
    for fn in (lib.qm_lattice_boundary_file, lib.qm_lattice_load_state, lib.qm_lattice_load_hfield):
        rc = fn(g.h, str(tmp_path / "missing.txt").encode())
        assert rc == -4, rc   # QM_EIO
        assert b"missing.txt" in lib.qm_last_error()
    bad = tmp_path / "outside_x0y0z0.txt"
    bad.write_text("3 4 20 0.1 0 0 0.1 0 0 0.5\n")
    assert lib.qm_lattice_load_state(g.h, str(bad).encode()) == -4
    cut = tmp_path / "cut.txt"
    cut.write_text("1\n0 5.3 0.53 3\n1 1 1 0 0 0 0 0\n")
    assert lib.qm_lattice_boundary_file(g.h, str(cut).encode()) == -4
    with pytest.raises(qmin.QminError):
        g.boundary_file(str(tmp_path / "missing.txt"))


@pytest.mark.gpu
def test_nan_state_ends_fire_with_qm_enan():
    """A NaN in Q makes forceMax NaN; the reference would report that run as
    converged (energyMinimizerFIRE.cpp:160: NaN > cutoff is false); the library
    stops the run and returns QM_ENAN, and the state reports nanDetected."""
    import qmin
    g = qmin.Lattice(16, 16, 16)
    g.random_init(3, S0)
    q = g.q()
    q[100, 2] = np.nan
    g.set_q(q)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=1)
    m = qmin.FIRE(max_iter=50)
    with pytest.raises(qmin.QminError, match="NaN"):
        m.minimize(f, g)
    st = m.state()
    assert st.nanDetected == 1 and st.converged == 0 and st.iterations == 1


@pytest.mark.gpu
def test_one_ensemble_lattice_with_nan_leaves_the_others_running():
    """In an ensemble the lattice whose force norm turns NaN stops at once; the
    others run to their own tolerance; the call returns QM_ENAN with every
    lattice's iterations and forceMax filled in."""
    import qmin
    lats = []
    for k in range(3):
        g = qmin.Lattice(20, 18, 16, precision=qmin.QM_FP32)
        g.random_init(10 + k, S0)
        g.sphere((10.0, 9.0, 8.0), 4.0, qmin.HOMEOTROPIC, 5.3, S0)
        lats.append(g)
    q = lats[1].q()
    q[7, 0] = np.nan
    lats[1].set_q(q)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=1)
    ens = qmin.Ensemble(lats)
    params = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=1e-3, max_iter=3000)
    lib = qmin.library()
    its = (ctypes.c_int * 3)()
    fms = (ctypes.c_double * 3)()
    rc = lib.qm_ensemble_fire_minimize(ens.h, params.h, f.h, 16, its, fms)
    assert rc == -6, lib.qm_last_error()
    assert its[1] == 1 and np.isnan(fms[1])
    for k in (0, 2):
        assert 1 < its[k] < 3000 and fms[k] <= 1e-3
        alone = qmin.Lattice(20, 18, 16, precision=qmin.QM_FP32)
        alone.random_init(10 + k, S0)
        alone.sphere((10.0, 9.0, 8.0), 4.0, qmin.HOMEOTROPIC, 5.3, S0)
        st = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=1e-3, max_iter=3000).minimize(f, alone)
        assert st.iterations == its[k]


@pytest.mark.gpu
def test_more_than_127_objects_match_oracle():
    """Object ids above 127 (site types are 16-bit on the device; the reference
    keeps int types): 150 small spheres of both anchorings and several W on a
    40^3 lattice.  Types, forces and energy against the oracle, then FIRE."""
    import qmin
    shape = (40, 40, 40)
    o = po.OracleSystem(*shape)
    g = qmin.Lattice(*shape)
    o.random_init(21, S0)
    g.random_init(21, S0)
    rng = np.random.default_rng(150)
    for k in range(150):
        c = tuple(float(v) for v in rng.uniform(0, 40, 3))
        bt, W = k % 2, float(rng.uniform(1, 10))
        o.sphere(c, 1.3, bt, W, S0)
        g.sphere(c, 1.3, bt, W, S0)
    t = g.types()
    assert np.array_equal(t, o.types())
    assert t.max() > 127, t.max()
    o.set_params(A, B, C, 2.32, 0.5, 0.3, multi=True, semantics=1)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=1)
    ref = o.forces()
    got = f.compute_forces(g)
    assert_forces_close(got, ref)
    e_ref, _ = o.energy()
    e_got, _ = f.compute_energy(g)
    assert abs(e_got - e_ref) <= 1e-12 * abs(e_ref)
    _, tr_ref = o.fire(max_iter=10)
    m = qmin.FIRE(max_iter=10)
    m.set_trace(16)
    m.minimize(f, g)
    np.testing.assert_allclose(m.trace()[:, 0], tr_ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(g.q(), o.q(), rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_lattice_on_a_caller_stream():
    """qm_lattice_set_stream: the lattice's work runs on a caller's stream (here
    torch's, default priority) and gives the same result as on its own; None
    returns to the lattice's own stream."""
    torch = pytest.importorskip("torch")
    import qmin
    if qmin._lib is not None and not qmin._torch_loaded_first:
        pytest.skip("libqmin.so was loaded before torch in this process (the -m gpu session imports torch first)")
    res = []
    for use in (False, True):
        g = qmin.Lattice(24, 20, 16)
        g.random_init(4, S0)
        g.sphere((12.0, 10.0, 8.0), 4.0, qmin.HOMEOTROPIC, 5.3, S0)
        s = torch.cuda.Stream()
        if use:
            g.set_stream(s.cuda_stream)
        f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=1)
        st = qmin.FIRE(max_iter=25).minimize(f, g)
        if use:
            g.set_stream(None)
            st2 = qmin.FIRE(max_iter=5).minimize(f, g)
            assert st2.iterations == 5
        res.append((st.forceMax, g.q()))
    assert res[0][0] == res[1][0]
