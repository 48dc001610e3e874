"""Extract the reference's L2/L3/L4/L6 force stencils as coefficient tables.

The reference's distortion forces (src/forces/lcForces.h:66-1043; dispatched
from src/forces/landauDeGennesLCDistortionTerms.cpp:113-330) are generated
polynomials in 125 per-site inputs:

    0..4      q at the site                       (qCurrent)
    5..34     Q of the 6 neighbours               (xDown, xUp, yDown, yUp, zDown, zUp)
    35..124   first-derivative vectors of those   (cubicLatticeDerivativeVector,
              neighbours, 15 each: d/dx q0..4,     layout dDimensionalVectorTypes.h:71-91)
              d/dy q0..4, d/dz q0..4

L2, L3 and L4 are homogeneous linear in these inputs and L6 is homogeneous
quadratic, in the bulk form and in each of the 27 boundary cases
(getBoundaryCase, lcForces.h:13-25).  This script runs the compiled reference
functions (oracle/_ref/libqmref.so, built from the reference sources by
oracle/Makefile) on unit vectors and pairs of unit vectors and records the
exact coefficients (all are dyadic rationals, so the probes are exact).  It
checks homogeneity and then re-evaluates every table on random inputs against
the reference to 1e-13.

Output: open-qmin_amd/stencils/ldg_stencil_tables.json (coefficients at L = 1).
Run here, where the reference checkout exists:  python tests/golden/extract_stencil_tables.py
"""
import ctypes
import json
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REFLIB = os.path.join(REPO, "oracle", "_ref", "libqmref.so")
OUT = os.path.join(REPO, "open-qmin_amd", "stencils", "ldg_stencil_tables.json")

NEIGHBOURS = ["xd", "xu", "yd", "yu", "zd", "zu"]
INPUT_NAMES = (["q%d" % i for i in range(5)]
               + ["%s.q%d" % (n, i) for n in NEIGHBOURS for i in range(5)]
               + ["%s.d%s%d" % (n, a, i) for n in NEIGHBOURS for a in "xyz" for i in range(5)])
# boundary case code = xx + 4*yy + 16*zz with 0 none, 1 up neighbour is an object,
# 2 down neighbour is an object (lcForces.h:13-25)
CASES = [xx + 4 * yy + 16 * zz for zz in range(3) for yy in range(3) for xx in range(3)]
LINEAR = (2, 3, 4)
QUADRATIC = (6,)


def load_ref():
    lib = ctypes.CDLL(REFLIB)
    D = ctypes.POINTER(ctypes.c_double)
    lib.qmref_lk.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, D, D, D, D]
    lib.qmref_lk.restype = None

    def lk(k, bcase, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(5)
        lib.qmref_lk(k, 1.0, bcase, x[:5].ctypes.data_as(D), x[5:35].ctypes.data_as(D),
                     x[35:125].ctypes.data_as(D), out.ctypes.data_as(D))
        return out
    return lk


def linear_table(lk, k, c):
    terms = []
    for i in range(125):
        e = np.zeros(125)
        e[i] = 1.0
        f = lk(k, c, e)
        for r in range(5):
            if f[r] != 0.0:
                terms.append([r, float(f[r]), i])
    return terms


def quadratic_table(lk, k, c):
    diag = np.zeros((125, 5))
    for i in range(125):
        e = np.zeros(125)
        e[i] = 1.0
        diag[i] = lk(k, c, e)
    terms = []
    for i in range(125):
        for r in range(5):
            if diag[i, r] != 0.0:
                terms.append([r, float(diag[i, r]), i, i])
        for j in range(i + 1, 125):
            e = np.zeros(125)
            e[i] = 1.0
            e[j] = 1.0
            f = lk(k, c, e) - diag[i] - diag[j]
            for r in range(5):
                if f[r] != 0.0:
                    terms.append([r, float(f[r]), i, j])
    return terms


def evaluate(terms, x):
    out = np.zeros(5)
    for t in terms:
        if len(t) == 3:
            out[t[0]] += t[1] * x[t[2]]
        else:
            out[t[0]] += t[1] * x[t[2]] * x[t[3]]
    return out


def main():
    if not os.path.exists(REFLIB):
        sys.exit("build oracle/_ref first: make -C oracle ref")
    lk = load_ref()
    rng = np.random.default_rng(20261202)
    tables = {}
    for k in LINEAR + QUADRATIC:
        per = {}
        for c in [-1] + CASES:
            x = rng.standard_normal(125)
            f1, f2, f0 = lk(k, c, x), lk(k, c, 2.0 * x), lk(k, c, np.zeros(125))
            power = 1.0 if k in LINEAR else 2.0
            assert not f0.any(), (k, c)
            assert np.allclose(f2, (2.0 ** power) * f1, rtol=1e-13, atol=1e-13), (k, c)
            terms = linear_table(lk, k, c) if k in LINEAR else quadratic_table(lk, k, c)
            for _ in range(4):
                x = rng.standard_normal(125)
                ref = lk(k, c, x)
                got = evaluate(terms, x)
                assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * max(1.0, np.abs(ref).max())), (k, c)
            per["bulk" if c < 0 else str(c)] = terms
        tables["L%d" % k] = per
        print("L%d: %d cases, bulk terms %d" % (k, len(per) - 1, len(per["bulk"])))
    doc = {
        "description": "Coefficients of the reference's per-site L2/L3/L4/L6 force stencils "
                       "(src/forces/lcForces.h), at L=1, measured by running the compiled reference "
                       "functions; entries are [component, coefficient, input] (linear) or "
                       "[component, coefficient, input_a, input_b] (quadratic).",
        "generator": "tests/golden/extract_stencil_tables.py",
        "inputs": INPUT_NAMES,
        "cases": CASES,
        "tables": tables,
    }
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=None, separators=(",", ":"))
        fh.write("\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
