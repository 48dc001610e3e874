"""Generate HIP device code for the L2/L3/L4/L6 force stencils.

Input: ldg_stencil_tables.json (per term and boundary case, entries
[component, coefficient, input] or [component, coefficient, input_a, input_b]
over the 125 per-site inputs: q[0..4], the 6 neighbours' Q [5..34] and their
first-derivative vectors [35..124], neighbour order xd, xu, yd, yu, zd, zu and
derivative order d/dx, d/dy, d/dz of q0..q4).

Output: ../csrc/gen/ldg_stencils.inc with

  ldg_boundary_lin(bcase, x, acc)         the 27 boundary-case L2/L3/L4 forms
  ldg_boundary_l6(bcase, x, out)          the 27 boundary-case L6 forms

acc[t][c] accumulates term t (0 = L2, 1 = L3, 2 = L4) of component c at unit
constant; the caller scales by L.

It also composes, for a site whose neighbours all have central derivatives,
the bulk L1 (shaped Laplacian), L2, L3 and L4 forms with those central
differences into one linear operator on the 19-point neighbourhood (centre,
6 faces, 12 edges).  Its weights are L1 C1 + L2 C2 + L3 C3 + L4 C4 per entry:

  ldg_fast_lin_a / _b(W, q, sm, sc, sp, h, acc)   device: acc[r] += W[k] Q_c(p)
  ../csrc/gen/ldg_fast_host.inc                   host: C1..C4 per entry k
  ldg_ring_lin_a / _b(W, q, sm, sc, sp, h, cm, cc, cp, acc)
                                      device: the same operator over rings of
                                      single planes, with the operator's few
                                      distinct weights (host: kPlaneLin*)
variant a covers L1..L3 (L4 = 0), variant b L1..L4.  The bulk L6 form, a
quadratic form in the same central differences, becomes one over the 19 point
values, generated together with variant b so that each value is read once:

  ldg_fast_lin_b_l6(W, q, sm, sc, sp, h, qc, acc, acc6)
                                      device: variant b into acc, and
                                      acc6[c] += x_a (sum_b coef x_b)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = os.path.join(HERE, "ldg_stencil_tables.json")
OUT = os.path.join(os.path.dirname(HERE), "csrc", "gen", "ldg_stencils.inc")
LIN_TERMS = ("L2", "L3", "L4")


def coef_literal(c):
    r = repr(float(c))
    if "e" not in r and "." not in r:
        r += ".0"
    return "R(%s)" % r


def xname(i):
    return "x[%d]" % i


def gen_l6(entries, fn_lines, indent="    "):
    acc = {c: [] for c in range(5)}
    for e in entries:
        acc[e[0]].append(e)
    for c in range(5):
        fn_lines.append("%s{" % indent)
        fn_lines.append("%s    R s = R(0);" % indent)
        for e in acc[c]:
            fn_lines.append("%s    s += %s * (%s * %s);" % (indent, coef_literal(e[1]), xname(e[2]), xname(e[3])))
        fn_lines.append("%s    out[%d] = s;" % (indent, c))
        fn_lines.append("%s}" % indent)


def gen_boundary(tables, cases):
    lines = ["template <typename R>",
             "__device__ __forceinline__ void ldg_boundary_lin(int bcase, const R* x, R (&acc)[3][5]) {",
             "    switch (bcase) {"]
    for c in cases:
        lines.append("    case %d: {" % c)
        for t, name in enumerate(LIN_TERMS):
            for e in tables[name][str(c)]:
                lines.append("        acc[%d][%d] += %s * %s;" % (t, e[0], coef_literal(e[1]), xname(e[2])))
        lines.append("        break;")
        lines.append("    }")
    lines += ["    default: break;", "    }", "}"]
    l6 = ["template <typename R>",
          "__device__ __forceinline__ void ldg_boundary_l6(int bcase, const R* x, R* out) {",
          "    switch (bcase) {"]
    for c in cases:
        l6.append("    case %d: {" % c)
        gen_l6(tables["L6"][str(c)], l6, "        ")
        l6.append("        break;")
        l6.append("    }")
    l6 += ["    default: for (int i = 0; i < 5; ++i) out[i] = R(0); break;", "    }", "}"]
    return lines + l6


POINTS = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
POINTS += [(a, b, 0) for a in (-1, 1) for b in (-1, 1)] + [(a, 0, b) for a in (-1, 1) for b in (-1, 1)]
POINTS += [(0, a, b) for a in (-1, 1) for b in (-1, 1)]
NBR = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
AXES = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]


def compose(entries):
    """Bulk table over (q, neighbour Q, neighbour derivatives) -> {(r, point, comp): coef}
    with the neighbour derivatives replaced by central differences."""
    from fractions import Fraction
    out = {}

    def add(r, p, c, v):
        key = (r, p, c)
        out[key] = out.get(key, Fraction(0)) + v

    for e in entries:
        r, coef, i = e[0], Fraction(e[1]), e[2]
        if i < 5:
            add(r, (0, 0, 0), i, coef)
        elif i < 35:
            d, c = (i - 5) // 5, (i - 5) % 5
            add(r, NBR[d], c, coef)
        else:
            j = i - 35
            d, a, c = j // 15, (j % 15) // 5, j % 5
            n, ax = NBR[d], AXES[a]
            up = tuple(n[k] + ax[k] for k in range(3))
            dn = tuple(n[k] - ax[k] for k in range(3))
            add(r, up, c, coef / 2)
            add(r, dn, c, -coef / 2)
    return {k: v for k, v in out.items() if v != 0}


def shaped_laplacian():
    """bulkL1Force at L1 = 1: s = 6 q - sum of faces, then s0 += s0 + s3,
    s3 += s0 + s3, s1, s2, s4 doubled."""
    from fractions import Fraction
    raw = {}
    for c in range(5):
        raw[(c, (0, 0, 0), c)] = Fraction(6)
        for p in NBR:
            raw[(c, p, c)] = Fraction(-1)
    shape = {0: {0: 2, 3: 1}, 1: {1: 2}, 2: {2: 2}, 3: {0: 1, 3: 2}, 4: {4: 2}}
    out = {}
    for (s, p, c), v in raw.items():
        for r, row in shape.items():
            if s in row:
                out[(r, p, c)] = out.get((r, p, c), Fraction(0)) + row[s] * v
    return {k: v for k, v in out.items() if v != 0}


# scheduling fence every 4 stencil points in the linear operators: bounds the
# values in flight (registers) without serialising the loads
FENCE_EVERY = 4


def gen_fast(tables):
    terms = [shaped_laplacian(), compose(tables["L2"]["bulk"]), compose(tables["L3"]["bulk"]),
             compose(tables["L4"]["bulk"])]
    dev, host = [], ["// Generated by open-qmin_amd/stencils/gen_stencils.py.  Do not edit by hand.",
                     "// Composed bulk operator: weight of entry k = L1 C[k][0] + L2 C[k][1] + L3 C[k][2] + L4 C[k][3].",
                     "#pragma once"]
    for name, used in (("a", 3), ("b", 4)):
        keys = sorted(set().union(*[set(t) for t in terms[:used]]),
                      key=lambda k: (POINTS.index(k[1]), k[2], k[0]))
        host.append("static const int kFastLin%sCount = %d;" % (name.upper(), len(keys)))
        host.append("static const double kFastLin%s[%d][4] = {" % (name.upper(), len(keys)))
        for k in keys:
            host.append("    {%s}," % ", ".join(repr(float(t.get(k, 0))) for t in terms))
        host.append("};")
        dev += ["template <int HX, int HP, typename R, typename WP>",
                "__device__ __forceinline__ void ldg_fast_lin_%s(WP W, const R (*q)[5][HP], int sm,"
                " int sc, int sp, int h, R* acc) {" % name,
                "    R v;"]
        last = None
        points_seen = []
        for idx, (r, p, c) in enumerate(keys):
            if p not in points_seen:
                # bound the number of stencil values in flight (register pressure)
                if points_seen and len(points_seen) % FENCE_EVERY == 0:
                    dev.append("    __builtin_amdgcn_sched_barrier(0);")
                points_seen.append(p)
            if (p, c) != last:
                slot = {-1: "sm", 0: "sc", 1: "sp"}[p[2]]
                off = "h"
                if p[0]:
                    off += " %s 1" % ("+" if p[0] > 0 else "-")
                if p[1]:
                    off += " %s HX" % ("+" if p[1] > 0 else "-")
                dev.append("    v = q[%s][%d][%s];" % (slot, c, off))
                last = (p, c)
            dev.append("    acc[%d] += W[%d] * v;" % (r, idx))
        dev.append("}")
        rows = [tuple(float(t.get(k, 0)) for t in terms) for k in keys]
        distinct = distinct_weights(rows)
        dev += gen_direct_distinct(name, keys, rows, distinct)
        host.append("// distinct weights of ldg_ring_lin_%s (each entry is one of them or its negative)" % name)
        host.append("static const int kPlaneLin%sCount = %d;" % (name.upper(), len(distinct)))
        host.append("static const double kPlaneLin%s[%d][4] = {" % (name.upper(), len(distinct)))
        for row in distinct:
            host.append("    {%s}," % ", ".join(repr(v) for v in row))
        host.append("};")
    return dev, host, keys


def distinct_weights(rows):
    """The distinct coefficient rows of an operator up to sign (first nonzero
    coefficient positive).  Negating a row negates its L-weighted sum exactly,
    so an entry may use W[j] or -W[j] in place of its own weight."""
    distinct = []
    for row in rows:
        lead = next(v for v in row if v != 0)
        canon = row if lead > 0 else tuple(-v for v in row)
        if canon not in distinct:
            distinct.append(canon)
    return distinct


def gen_direct_distinct(name, keys, rows, distinct):
    """ldg_fast_lin_<name> over a ring of three planes of single-plane rows
    (q[slot][c][NP], row stride HXS), with the distinct weights (a table per
    variant, emitted for the host); the site's own column (the points (0, 0, -1), (0, 0, 0)
    and (0, 0, 1)) comes from registers cm, cc, cp instead of the ring:
    ldg_ring_lin_<name><HXS, NP, SB>(W, q, sm, sc, sp, h, cm, cc, cp, acc): a
    scheduling barrier after every SB points (0: none) bounds how many of the
    ring's values the compiler keeps in flight."""
    lines = ["template <int HXS, int NP, int SB, typename R, typename WP>",
             "__device__ __forceinline__ void ldg_ring_lin_%s(WP W, const R (*q)[5][NP], int sm, int sc, int sp,"
             " int h, const R* cm, const R* cc, const R* cp, R* acc) {" % name,
             "    R v;"]
    own = {(0, 0, -1): "cm", (0, 0, 0): "cc", (0, 0, 1): "cp"}
    last = None
    points_seen = []
    for (r, p, c), row in zip(keys, rows):
        if p not in points_seen:
            if points_seen:
                lines.append("    if constexpr (SB > 0 && %d %% SB == 0) __builtin_amdgcn_sched_barrier(0);"
                             % len(points_seen))
            points_seen.append(p)
        if (p, c) != last:
            if p in own:
                lines.append("    v = %s[%d];" % (own[p], c))
            else:
                slot = {-1: "sm", 0: "sc", 1: "sp"}[p[2]]
                off = "h"
                if p[0]:
                    off += " %s 1" % ("+" if p[0] > 0 else "-")
                if p[1]:
                    off += " %s HXS" % ("+" if p[1] > 0 else "-")
                lines.append("    v = q[%s][%d][%s];" % (slot, c, off))
            last = (p, c)
        lead = next(v for v in row if v != 0)
        canon = row if lead > 0 else tuple(-v for v in row)
        lines.append("    acc[%d] %s= W[%d] * v;" % (r, "+" if lead > 0 else "-", distinct.index(canon)))
    lines.append("}")
    return lines


L6_FACES = ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))


def l6_point_form(k):
    """Stencil input k as a linear form over the 19-point neighbourhood:
    {(point, component): weight}; a neighbour's transverse derivative is the
    central difference 0.5 (Q(up) - Q(down)) of two edge points."""
    if k < 5:
        return {((0, 0, 0), k): 1.0}
    if k < 35:
        d, c = divmod(k - 5, 5)
        return {(L6_FACES[d], c): 1.0}
    d, rest = divmod(k - 35, 15)
    ax, c = divmod(rest, 5)
    assert ax != d // 2, "along-axis neighbour derivatives are not inputs"
    up, dn = list(L6_FACES[d]), list(L6_FACES[d])
    up[ax] += 1
    dn[ax] -= 1
    return {(tuple(up), c): 0.5, (tuple(dn), c): -0.5}


def l6_point_kind(p):
    return ("centre", "face", "edge")[sum(abs(v) for v in p)]


def l6_lds(p, c):
    slot = {-1: "sm", 0: "sc", 1: "sp"}[p[2]]
    off = "h"
    if p[0]:
        off += " %s 1" % ("+" if p[0] > 0 else "-")
    if p[1]:
        off += " %s HX" % ("+" if p[1] > 0 else "-")
    return "q[%s][%d][%s]" % (slot, c, off)


def gen_fast_l6(entries, lin_keys):
    """Bulk L1..L4 operator (variant b, weights W) and L6 form in one pass over
    the 19 point values.  The L6 form is a quadratic form over those values;
    every product pairs a face value with the centre, another face or an edge
    value (the generator checks that), so the 30 face values are read from LDS
    once into registers and each edge point's 5 values once as the loop over
    edge points reaches it, and each value read also feeds its linear terms:
    90 LDS reads per site for both forms.  L6 products sharing a first factor
    x_a are summed as acc6[c] += x_a * (sum_b coef x_b)."""
    pair = {}
    for comp, coef, a, b in entries:
        for ka, wa in l6_point_form(a).items():
            for kb, wb in l6_point_form(b).items():
                key = (comp,) + tuple(sorted((ka, kb), key=lambda t: (l6_point_kind(t[0]) != "edge", t)))
                pair[key] = pair.get(key, 0.0) + coef * wa * wb
    pair = {k: v for k, v in pair.items() if v != 0.0}
    blocks = {}          # None: centre/face products; edge point: its products with faces
    for (comp, ka, kb), coef in pair.items():
        kinds = sorted((l6_point_kind(ka[0]), l6_point_kind(kb[0])))
        assert kinds in (["centre", "face"], ["face", "face"], ["edge", "face"]), kinds
        edge = ka[0] if kinds[0] == "edge" else None
        blocks.setdefault(edge, []).append((comp, ka, kb, coef))

    def name(k):
        p, c = k
        kind = l6_point_kind(p)
        if kind == "centre":
            return "qc[%d]" % c
        if kind == "face":
            return "f%d_%d" % (L6_FACES.index(p), c)
        return "e%d" % c

    def emit_block(terms, indent):
        # greedy grouping: the (first factor, component) pair covering the most
        # products not yet placed becomes the next group (one multiply-add each)
        left = list(terms)
        out = []
        while left:
            count = {}
            for comp, ka, kb, coef in left:
                for k in (ka, kb):
                    count[(k, comp)] = count.get((k, comp), 0) + 1
            best = max(sorted(count), key=lambda g: count[g])
            k, comp = best
            group = [(coef, kb if ka == k else ka) for c2, ka, kb, coef in left if c2 == comp and k in (ka, kb)]
            left = [t for t in left if not (t[0] == comp and k in (t[1], t[2]))]
            out.append("%s{" % indent)
            coef, b = group[0]
            out.append("%s    R t = %s * %s;" % (indent, coef_literal(coef), name(b)))
            for coef, b in group[1:]:
                out.append("%s    t += %s * %s;" % (indent, coef_literal(coef), name(b)))
            out.append("%s    acc6[%d] += %s * t;" % (indent, comp, name(k)))
            out.append("%s}" % indent)
        return out

    lin_at = {}
    for idx, (r, p, c) in enumerate(lin_keys):
        lin_at.setdefault((p, c), []).append((idx, r))

    def emit_lin(p, c, value, indent):
        return ["%sacc[%d] += W[%d] * %s;" % (indent, r, idx, value) for idx, r in lin_at.pop((p, c), [])]

    lines = ["template <int HX, int HP, typename R, typename WP>",
             "__device__ __forceinline__ void ldg_fast_lin_b_l6(WP W, const R (*q)[5][HP], int sm, int sc, int sp,"
             " int h, const R* qc, R* acc, R* acc6) {"]
    for c in range(5):
        lines += emit_lin((0, 0, 0), c, "qc[%d]" % c, "    ")
    for d, p in enumerate(L6_FACES):
        for c in range(5):
            lines.append("    const R f%d_%d = %s;" % (d, c, l6_lds(p, c)))
            lines += emit_lin(p, c, "f%d_%d" % (d, c), "    ")
    lines += emit_block(sorted(blocks[None]), "    ")
    edge_pts = sorted({p for (p, c) in lin_at} | {k for k in blocks if k is not None})
    for p in edge_pts:
        # one edge point per block: the fence keeps the compiler from hoisting
        # the loads of later edge points (and their registers) above it
        lines.append("    __builtin_amdgcn_sched_barrier(0);")
        lines.append("    {")
        for c in range(5):
            lines.append("        const R e%d = %s;" % (c, l6_lds(p, c)))
            lines += emit_lin(p, c, "e%d" % c, "        ")
        lines += emit_block(sorted(blocks.get(p, [])), "        ")
        lines.append("    }")
    assert not lin_at, "linear terms outside the 19-point neighbourhood"
    lines.append("}")
    return lines


def main():
    with open(TABLES) as fh:
        doc = json.load(fh)
    tables, cases = doc["tables"], doc["cases"]
    body = ["// Generated by open-qmin_amd/stencils/gen_stencils.py from ldg_stencil_tables.json.",
            "// Do not edit by hand.", "#pragma once", ""]
    body += gen_boundary(tables, cases)
    fast_dev, fast_host, lin_keys = gen_fast(tables)
    body += [""] + fast_dev
    body += [""] + gen_fast_l6(tables["L6"]["bulk"], lin_keys)
    host_text = "\n".join(fast_host) + "\n"
    host_out = os.path.join(os.path.dirname(OUT), "ldg_fast_host.inc")
    if not os.path.exists(host_out) or open(host_out).read() != host_text:
        with open(host_out, "w") as fh:
            fh.write(host_text)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    text = "\n".join(body) + "\n"
    if not os.path.exists(OUT) or open(OUT).read() != text:
        with open(OUT, "w") as fh:
            fh.write(text)
    print("wrote", OUT, len(text), "bytes")


if __name__ == "__main__":
    main()
