"""The reference's own openQmin driver body, unchanged, over the drop-in C++
classes.

oracle/Makefile (target `driver`) copies the body of openQmin.cpp's main() (from
the phase constants to the class-size report: the lattice, force, fields, FIRE
updater, setInitialConditions.h, boundary file, addObjectsToOpenQmin.h, the
three save schedules, energy and the size report) out of the reference checkout
and compiles it against open-qmin_amd/host/openqmin_hip.h and libqmin;
oracle/ref_driver_main.cpp stands where the MPI/TCLAP preamble was.  Building it
needs the checkout (this container); the binary lands in oracle/_ref and runs
on the GPU box, where the run is compared with the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from conftest import REFERENCE_CHECKOUT, REPO

BODY = os.path.join(REPO, "oracle", "_ref", "openQmin-body")
A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE_CHECKOUT, "openQmin.cpp")),
                    reason="needs the reference checkout")
def test_reference_driver_body_compiles_against_the_drop_in_classes():
    lib = os.path.join(REPO, "open-qmin_amd", "lib", "libqmin.so")
    if not os.path.exists(lib):
        pytest.skip("libqmin.so not built")
    p = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "driver"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert p.returncode == 0, p.stdout
    assert os.path.exists(BODY)
    body = open(os.path.join(REPO, "oracle", "_ref", "driver_body.inc")).read()
    ref = open(os.path.join(REFERENCE_CHECKOUT, "openQmin.cpp")).read()
    assert body in ref and body.count("\n") > 150


def _final_line(out):
    line = [l for l in out.splitlines() if l.startswith("minimized to")][-1]
    fm, e = line.split()[2], line.split("E=")[1]
    return float(fm), float(e)


@pytest.mark.gpu
@pytest.mark.parametrize("gpu_flag", [False, True])
def test_reference_driver_body_runs_and_matches_oracle(tmp_path, gpu_flag):
    """`openQmin -l 20 -i 40 --L1 2.32 --L2 0.5 --randomSeed 1234 --saveFile ...`
    through the reference's own driver body: without -g the body selects the CPU
    path's semantics (sim->setCPUOperation(!GPU)), with -g 0 the GPU path's.
    Final forceMax and energy (printed with %g and %f) and the saved state (six
    decimals) against the oracle's FIRE run from the same seed."""
    if not os.path.exists(BODY):
        pytest.skip("oracle/_ref/openQmin-body not built (needs the reference checkout)")
    save = str(tmp_path / "final")
    args = [BODY, "-l", "20", "-i", "40", "--L1", "2.32", "--L2", "0.5", "--randomSeed", "1234", "--saveFile", save]
    if gpu_flag:
        args += ["-g", "0"]
    print("reference driver body:", " ".join(args[1:]), flush=True)
    p = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    fm, e = _final_line(p.stdout)
    o = po.OracleSystem(20, 20, 20)
    o.random_init(1234, S0)
    o.set_params(A, B, C, 2.32, 0.5, 0, 0, 0, multi=True, semantics=1 if gpu_flag else 0)
    _, tr = o.fire(max_iter=40)
    e_ref, _ = o.energy()
    assert fm == pytest.approx(tr[-1, 0], rel=1e-5)
    assert e == pytest.approx(e_ref / o.n_active(), abs=1e-6)
    rows = np.loadtxt(save + "_x0y0z0.txt")
    site = (rows[:, 0] + 20 * (rows[:, 1] + 20 * rows[:, 2])).astype(int)
    q = np.empty((site.size, 5))
    q[site] = rows[:, 3:8]
    np.testing.assert_allclose(q, o.q(), atol=1e-6)
    assert "size of configuration" not in p.stdout   # the report is verbose-only
