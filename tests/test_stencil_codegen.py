"""The generated bulk operator with L6 (csrc/gen/ldg_stencils.inc,
ldg_fast_lin_b_l6) against the bulk L6 table it is generated from, and its
linear part against the separately generated L1..L4 operator (ldg_fast_lin_b).

The table (ldg_stencil_tables.json, "L6"/"bulk") lists the products of the
reference's bulkL6Force (src/forces/lcForces.h:801-816) over its 125 inputs:
the site's Q, the six face neighbours' Q and their first-derivative vectors,
which for a bulk neighbour are central differences of the two edge points
beside it (landauDeGennesLCOtherForces.cpp:15-145).  The device operator is a
regrouped quadratic form over the 19 point values read from the LDS ring,
generated into one function with the composed L1..L4 operator.  Here the text
of the generated functions is executed as Python on random point values and
weights and compared with the table form evaluated from the same values, so a
dropped product, a wrong coefficient, a wrong edge point or a mis-signed
central difference in the generator fails.
"""
import json
import os
import re

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLES = os.path.join(REPO, "open-qmin_amd", "stencils", "ldg_stencil_tables.json")
GENERATED = os.path.join(REPO, "open-qmin_amd", "csrc", "gen", "ldg_stencils.inc")

FACES = ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))
HX = 3          # a 3 x 3 plane per slot, x fastest
H = 4           # the centre of that plane
SLOT = {"sm": 0, "sc": 1, "sp": 2}


def table_form(entries, P):
    """Bulk L6 table evaluated as the reference's inputs define it.  P[(x, y, z)]
    holds the 5 components of point (x, y, z) of the 19-point neighbourhood."""
    x = np.zeros(125)
    x[0:5] = P[(0, 0, 0)]
    for d, f in enumerate(FACES):
        x[5 + 5 * d:10 + 5 * d] = P[f]
        for ax in range(3):
            if ax == d // 2:
                continue
            up, dn = list(f), list(f)
            up[ax] += 1
            dn[ax] -= 1
            x[35 + 15 * d + 5 * ax:40 + 15 * d + 5 * ax] = 0.5 * (P[tuple(up)] - P[tuple(dn)])
    out = np.zeros(5)
    for comp, coef, a, b in entries:
        out[comp] += coef * x[a] * x[b]
    return out


def generated_body(name="ldg_fast_lin_b_l6"):
    text = open(GENERATED).read()
    start = text.index("void %s(" % name)
    start = text.index("{", start) + 1
    end = text.index("\n}\n", start)
    return text[start:end]


def run_generated(body, P, W=None):
    """Execute the generated C++ statements as Python over an LDS ring
    q[slot][c][pos]; returns (acc, acc6)."""
    q = np.zeros((3, 5, 9))
    for p, vals in P.items():
        q[p[2] + 1, :, H + p[0] + HX * p[1]] = vals
    env = {"q": q, "qc": P[(0, 0, 0)].copy(), "acc": np.zeros(5), "acc6": np.zeros(5), "h": H, "HX": HX,
           "R": float, "W": np.zeros(200) if W is None else W}
    env.update(SLOT)
    code = []
    for line in body.split("\n"):
        s = line.strip()
        if not s or s in ("{", "}") or s.startswith("__builtin_amdgcn_sched_barrier"):
            continue
        assert s.endswith(";"), s
        s = s[:-1]
        s = re.sub(r"^(const )?R ", "", s)
        if "=" not in s:
            continue                        # a bare declaration
        code.append(s)
    exec("\n".join(code), env)
    return env["acc"], env["acc6"]


def points(rng):
    P = {(0, 0, 0): rng.standard_normal(5)}
    for f in FACES:
        P[f] = rng.standard_normal(5)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for sa in (-1, 1):
            for sb in (-1, 1):
                p = [0, 0, 0]
                p[a], p[b] = sa, sb
                P[tuple(p)] = rng.standard_normal(5)
    return P


def test_generated_l6_matches_table_form():
    entries = json.load(open(TABLES))["tables"]["L6"]["bulk"]
    body = generated_body()
    rng = np.random.default_rng(20261202)
    for _ in range(20):
        P = points(rng)
        want = table_form(entries, P)
        acc, got = run_generated(body, P)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
        assert not acc.any()


def test_linear_part_matches_linear_operator():
    body, lin = generated_body(), generated_body("ldg_fast_lin_b")
    rng = np.random.default_rng(99)
    for _ in range(5):
        P = points(rng)
        W = rng.standard_normal(200)
        acc, _ = run_generated(body, P, W)
        want, none = run_generated(lin, P, W)
        assert not none.any()
        np.testing.assert_allclose(acc, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_generated_l6_reads_each_point_once():
    # 30 face values and 60 edge values, each read from LDS once per site (the
    # centre comes in registers)
    reads = re.findall(r"= q\[(s[mcp])\]\[(\d)\]\[([^\]]+)\];", generated_body())
    assert len(reads) == 90
    assert len(set(reads)) == 90


def test_generated_l6_is_not_trivially_zero():
    # one component's quadratic form is not identically zero along a single point's values
    entries = json.load(open(TABLES))["tables"]["L6"]["bulk"]
    rng = np.random.default_rng(7)
    P = points(rng)
    P2 = {k: v.copy() for k, v in P.items()}
    P2[(1, 1, 0)] += 1.0
    assert np.abs(table_form(entries, P2) - table_form(entries, P)).max() > 1e-3
    assert np.abs(run_generated(generated_body(), P2)[1] - run_generated(generated_body(), P)[1]).max() > 1e-3
