"""The C++ drop-in classes (open-qmin_amd/host/openqmin_hip.h) through the
openQmin-style driver built on them, against the CPU oracle.

The driver is C++ host code calling libqmin through the C ABI only, as a
reference driver would after switching its includes.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "open-qmin_amd", "lib", "openQmin-hip")

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)


def run(*args, check=True):
    p = subprocess.run([DRIVER] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    if check:
        assert p.returncode == 0, p.stderr
    return p


def summary(p):
    toks = p.stdout.strip().splitlines()[-1].split()
    d = dict(zip(toks[0::2], toks[1::2]))
    return int(d["iterations"]), float(d["forceMax"]), float(d["energy"])


def test_driver_rejects_unknown_and_unsupported_flags():
    assert os.path.exists(DRIVER), "build the package first"
    for args in (["--bogus", "3"], ["-z", "4"], ["-z", "-1"], ["--precision", "mixed", "--linearSpacedSaving", "5"],
                 ["--wall", "2,0"], ["--precision", "fp16"]):
        p = run(*args, check=False)
        assert p.returncode == 1
        assert "openQmin-hip:" in p.stderr and "ROCm" not in p.stderr, p.stderr


@pytest.mark.gpu
def test_driver_fire_matches_oracle_one_constant():
    """openQmin -i 60 -l 24 --L1 2.32 --randomSeed 1234 (config-1 physics, smaller box)."""
    p = run("-l", 24, "-i", 60, "--L1", 2.32, "--randomSeed", 1234)
    it, fm, e = summary(p)
    o = po.OracleSystem(24, 24, 24)
    o.random_init(1234, S0)
    o.set_params(A, B, C, 2.32, 0, 0, semantics=1)
    it_ref, tr = o.fire(max_iter=60)
    e_ref, _ = o.energy()
    assert it == it_ref == 60
    assert fm == pytest.approx(tr[-1, 0], rel=1e-10)
    assert e == pytest.approx(e_ref / o.n_active(), rel=1e-12)


@pytest.mark.gpu
def test_driver_objects_multi_constant_and_saved_state(tmp_path):
    save = tmp_path / "final"
    p = run("--Lx", 22, "--Ly", 18, "--Lz", 16, "-i", 40, "--L1", 2.32, "--L2", 0.5, "--L3", 0.3, "--randomSeed",
            77, "--wall", "2,0,1,5.3", "--wall", "2,15,1,5.3", "--sphere", "11,9,8,4.5,0,5.3", "--cpuSemantics",
            "--saveFile", save)
    it, fm, e = summary(p)
    o = po.OracleSystem(22, 18, 16)
    o.random_init(77, S0)
    o.wall(2, 0, 1, 5.3, S0)
    o.wall(2, 15, 1, 5.3, S0)
    o.sphere((11, 9, 8), 4.5, 0, 5.3, S0)
    o.set_params(A, B, C, 2.32, 0.5, 0.3, semantics=0)
    _, tr = o.fire(max_iter=40)
    e_ref, _ = o.energy()
    assert fm == pytest.approx(tr[-1, 0], rel=1e-10)
    assert e == pytest.approx(e_ref / o.n_active(), rel=1e-12)
    rows = np.loadtxt(str(save) + "_x0y0z0.txt")
    assert rows.shape == (22 * 18 * 16, 10)
    site = (rows[:, 0] + 22 * (rows[:, 1] + 18 * rows[:, 2])).astype(int)
    q = np.empty((site.size, 5))
    q[site] = rows[:, 3:8]
    np.testing.assert_allclose(q, o.q(), atol=1e-6)   # the state file holds 6 decimals (%f)
    t = np.empty(site.size, dtype=int)
    t[site] = rows[:, 8]
    assert np.array_equal(t, o.types())


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["adam", "nesterov", "lolbfgs"])
def test_driver_other_minimizers(name):
    p = run("-l", 16, "-i", 25, "--L1", 2.32, "--L2", 0.5, "--randomSeed", 5, "-e", 1e-3, "--minimizer", name,
            "--sphere", "8,8,8,3.5,0,5.3")
    it, fm, _ = summary(p)
    o = po.OracleSystem(16, 16, 16)
    o.random_init(5, S0)
    o.sphere((8, 8, 8), 3.5, 0, 5.3, S0)
    o.set_params(A, B, C, 2.32, 0.5, 0, semantics=1)
    if name == "adam":
        _, tr = o.adam(dt=1e-3, fc=1e-12, max_iter=25)
    elif name == "nesterov":
        _, tr = o.nesterov(dt=1e-3, mu=0.01, fc=1e-12, max_iter=25)
    else:
        _, tr = o.lolbfgs(m=5, dt=1e-3, fc=1e-12, max_iter=25)
    assert it == 25
    assert fm == pytest.approx(tr[-1], rel=1e-9)


@pytest.mark.gpu
def test_driver_mixed_precision_polish_reaches_fp64_answer():
    """FP32 FIRE to the tolerance, then the FP64 polish: same converged energy as
    an all-FP64 run to 1e-9 relative, at a tolerance tight enough that the
    energy is settled."""
    common = ["-l", 20, "-i", 20000, "--L1", 2.32, "--L2", 0.5, "--randomSeed", 3, "-f", 1e-11,
              "--sphere", "10,10,10,4,0,5.3"]
    it64, fm64, e64 = summary(run(*common, "--precision", "fp64"))
    itm, fmm, em = summary(run(*common, "--precision", "mixed"))
    assert fm64 <= 1e-11 and fmm <= 1e-11
    assert em == pytest.approx(e64, rel=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("z", [1, 2, 3])
def test_driver_initial_condition_modes(z):
    """-z 1 (one random director from the seeded stream), -z 2 (the fixed
    director (1/2, 0, 1/2)), -z 3 (the example director function of
    setInitialConditions.h), then a homeotropic sphere (a uniform texture alone
    is already stationary) and 20 FIRE steps, against the oracle started the
    same way."""
    from conftest import example_directors
    import math
    p = run("-l", 16, "-i", 20, "--L1", 2.32, "--randomSeed", 99, "-z", z, "--sphere", "8,8,8,4,0,5.3")
    it, fm, e = summary(p)
    o = po.OracleSystem(16, 16, 16)
    if z == 1:
        o.random_uniform_init(99, S0)
    elif z == 2:
        a = math.sqrt(0.5 * 0.5 + 0.0 + 0.5 * 0.5)
        o.director_init(np.tile([0.5 / a, 0.0, 0.5 / a, S0], (o.n, 1)))
    else:
        o.director_init(example_directors((16, 16, 16)))
    o.sphere((8.0, 8.0, 8.0), 4.0, 0, 5.3, S0)
    o.set_params(A, B, C, 2.32, 0, 0, semantics=1)
    it_ref, tr = o.fire(max_iter=20)
    e_ref, _ = o.energy()
    assert it == it_ref == 20
    assert fm == pytest.approx(tr[-1, 0], rel=1e-10)
    assert e == pytest.approx(e_ref / o.n_active(), rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [("--linearSpacedSaving", "7"), ("--logSpacedSaving", "0.3")])
def test_driver_save_schedules(tmp_path, schedule):
    """openQmin.cpp:291-333: staged minimization with a state saved before each
    stage, named <saveFile>_t<iteration>; the stages follow the reference's loop
    (linear: while it + k < max; log: round(10^(0.3 j)) = 2, 4, 8, 16, 32) and
    FIRE continues across them as one run; --stride thins the saved sites."""
    base = str(tmp_path / "st")
    p = run("-l", 12, "-i", 30, "--L1", 2.32, "--randomSeed", 5, "--saveFile", base, "--stride", 2, *schedule)
    it, fm, e = summary(p)
    saved = sorted(int(f.name.split("_t")[1].split("_")[0]) for f in tmp_path.glob("st_t*_x0y0z0.txt"))
    if schedule[0] == "--linearSpacedSaving":
        assert saved == [0, 7, 14, 21] and it == 28
    else:
        assert saved == [0, 2, 4, 8, 16] and it == 32
    o = po.OracleSystem(12, 12, 12)
    o.random_init(5, S0)
    o.set_params(A, B, C, 2.32, 0, 0, semantics=1)
    it_ref, tr = o.fire(max_iter=it)
    assert fm == pytest.approx(tr[-1, 0], rel=1e-10)
    rows = np.loadtxt(str(tmp_path / "st_x0y0z0.txt"))
    assert rows.shape == (6 ** 3, 10)
    assert np.all(rows[:, :3] % 2 == 0)
    q = o.q().reshape(12, 12, 12, 5)[::2, ::2, ::2].reshape(-1, 5)
    np.testing.assert_allclose(rows[:, 3:8], q, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_reference_array_views_and_force_into_array():
    """open-qmin_amd/host/array_views_example.cpp: reference-style code that
    edits Q through a read-write ArrayHandle on returnPositions(), reads the
    types through returnTypes(), and calls computeForces(GPUArray<dVec>&, zero,
    type) into the model's array and into its own; forces against the oracle
    on the same state, the added call doubling them exactly, a device handle
    refused."""
    exe = os.path.join(REPO, "open-qmin_amd", "lib", "array-views-example")
    p = subprocess.run([exe, "12", "77"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    out = {}
    for line in p.stdout.splitlines():
        key, *vals = line.split()
        out.setdefault(key, []).append(vals)
    o = po.OracleSystem(12, 12, 12)
    o.random_init(77, S0)
    o.sphere((6.0, 6.0, 6.0), 3.0, 0, 5.3, S0)
    q = o.q()
    q[1, 0] += 0.125
    q[1, 4] -= 0.0625
    o.set_q(q)
    assert [float(v) for v in out["q1"][0]] == [q[1, 0], q[1, 4]]
    assert int(out["objects"][0][0]) == int((o.types() > 0).sum())
    assert float(out["doubled"][0][0]) == 0.0
    o.set_params(A, B, C, 2.32, 0.5, 0.3, multi=True, semantics=1)
    ref = o.forces()
    for vals in out["f"]:
        i = int(vals[0])
        np.testing.assert_allclose([float(v) for v in vals[1:]], ref[i], rtol=1e-12, atol=1e-13)
    assert out["device-handle-throws"][0][0] == "1"
