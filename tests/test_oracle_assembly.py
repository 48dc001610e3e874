"""Pins of the oracle's whole-lattice force assembly that do not go through the
oracle itself.

The per-site pieces are the reference's compiled functions (oracle/_ref: the
stencils of src/forces/lcForces.h, getBoundaryCase, the anchoring pair force of
landauDeGennesLCBoundary.h, the phase derivatives of inc/qTensorFunctions.h, the
neighbour table of cubicLattice).  How the reference wires them together is
read here from its source text at test time:

* which neighbour feeds which stencil argument, and which neighbour's
  derivative vector (landauDeGennesLCDistortionTerms.cpp:16-330);
* the argument order of getBoundaryCase and boundaryL1Force's type arguments;
* which object neighbours contribute an anchoring pair force and with whose Q
  (landauDeGennesLCOtherForces.cpp:314-373);
* the uniform field force components (landauDeGennesLCOtherForces.cpp:285-312)
  and the field energy (landauDeGennesLC.cpp:238-253);
* the metric correction (landauDeGennesLC.cpp:137-160).

The control flow around them (which pass handles which site type, the two
metric corrections of the CPU path, multirankSimulation.cpp:452-473 and
landauDeGennesLC.cpp:125-202) is restated below from reading.  Finite
differences then pin the anchoring and field forces against their energies.
The tests need the reference checkout and oracle/_ref, i.e. they run in the
build container only.
"""
import os
import re
import types

import numpy as np
import pytest

import pyoracle as po
from conftest import REFERENCE_CHECKOUT

FORCES = os.path.join(REFERENCE_CHECKOUT, "src", "forces")
pytestmark = pytest.mark.skipif(not (po.ref_available() and os.path.isdir(FORCES)),
                                reason="needs the reference checkout and oracle/_ref")

A, B, C = po.rescaled_phase()
S0 = po.uniaxial_s0(A, B, C)
SHAPE = (9, 8, 7)
# objects in creation order: (kind, btype, W); walls at z = 0 and z = 6, spheres
# (center, radius)
OBJECTS = (("wall", 0, 1, 5.3), ("wall", 6, 0, 3.1), ("sphere", ((3.0, 3.5, 3.0), 1.6), 0, 2.2),
           ("sphere", ((6.5, 5.0, 3.5), 1.5), 1, 1.7))
E = dict(field=(0.3, -0.2, 0.5), eps=1.2, eps0=0.9, deps=0.7)
H = dict(field=(-0.4, 0.1, 0.6), chi=1.1, mu0=0.8, dchi=0.5)


def _src(name):
    with open(os.path.join(FORCES, name)) as fh:
        return fh.read()


def _body(text, head):
    """Text of the function body that follows `head` (brace matched)."""
    i = text.index(head)
    j = text.index("{", i)
    depth = 0
    for k in range(j, len(text)):
        depth += {"{": 1, "}": -1}.get(text[k], 0)
        if depth == 0:
            return text[j:k + 1]
    raise AssertionError("unbalanced braces after " + head)


def _args(call):
    return [a.strip() for a in call.split(",")]


def _wiring(body):
    """name -> neighbour slot (0..5) of every dVec and derivative vector the
    function body reads, as assigned in its text."""
    idx = {v: int(d) for v, d in re.findall(
        r"\b(i[xyz][du])\s*=\s*latticeNeighbors\.data\[lattice->neighborIndex\((\d),currentIndex\)\]", body)}
    assert sorted(idx.values()) == list(range(6)), idx
    q = {name: idx[v] for name, v in re.findall(r"\b([xyz](?:Down|Up))\s*=\s*Qtensors\.data\[(i[xyz][du])\]", body)}
    d = {name: idx[v] for name, v in
         re.findall(r"\b([xyz](?:Down|Up)Derivative)\s*=\s*h_derivatives\.data\[(i[xyz][du])\]", body)}
    t = {"latticeTypes.data[%s]" % v: s for v, s in idx.items()}
    return q, d, t


def _signature(lc, name):
    m = re.search(r"HOSTDEVICE\s+\w+\s+%s\((.*?)\)\s*\{" % name, lc, re.S)
    assert m, name
    return [a.split()[-1].lstrip("&") for a in _args(m.group(1))]


class Reference:
    """The reference's force assembly, from its compiled pieces and the
    wiring read from its source."""

    def __init__(self):
        self.lib = po.ref_lib()
        lc = _src("lcForces.h")
        dist = _src("landauDeGennesLCDistortionTerms.cpp")
        other = _src("landauDeGennesLCOtherForces.cpp")
        main = _src("landauDeGennesLC.cpp")
        # the harness (oracle/ref_harness.cpp) passes each argument in the
        # function's own parameter order: check that order here
        nb = ["xDown", "xUp", "yDown", "yUp", "zDown", "zUp"]
        nd = [x + "Derivative" for x in nb]
        for k in (2, 3, 4, 6):
            assert _signature(lc, "bulkL%dForce" % k) == ["L%d" % k, "qCurrent"] + nb + nd + ["spatialTerm"]
            assert _signature(lc, "boundaryL%dForce" % k) == (["L%d" % k, "boundaryCase", "qCurrent"] + nb + nd +
                                                             ["spatialTerm"])
        assert _signature(lc, "bulkL1Force") == ["L1", "qCurrent"] + nb + ["spatialTerm"]
        assert _signature(lc, "boundaryL1Force")[:8] == ["L1", "qCurrent"] + nb
        assert _signature(lc, "getBoundaryCase") == ["xd", "xu", "yd", "yu", "zd", "zu"]
        self.calls = {}
        for fn in ("computeL1BulkCPU", "computeL1BoundaryCPU", "computeAllDistortionTermsBulkCPU",
                   "computeAllDistortionTermsBoundaryCPU"):
            body = _body(dist, "void landauDeGennesLC::%s(" % fn)
            q, d, t = _wiring(body)
            assert "force -= a*derivativeTrQ2(qCurrent);" in body
            assert "force -= b*derivativeTrQ3(qCurrent);" in body
            assert "force -= c*derivativeTrQ2Squared(qCurrent);" in body
            assert re.search(r"scalar a = 0\.5\*A;\s*scalar b = B/3\.0;\s*scalar c = 0\.25\*C;", body)
            calls = []
            for kind, k, a in re.findall(r"lcForce::(bulk|boundary)L(\d)Force\((.*?)\);", body, re.S):
                a = _args(a)
                entry = {"kind": kind, "k": int(k)}
                rest = a[2:] if kind == "bulk" or k == "1" else a[3:]
                entry["nb"] = [q[x] for x in rest[:6]]
                if k == "1" and kind == "boundary":
                    entry["types"] = [t[x] for x in rest[6:12]]
                elif k != "1":
                    entry["nd"] = [d[x] for x in rest[6:12]]
                    if kind == "boundary":
                        assert a[1] == "boundaryCase"
                calls.append(entry)
            if "Boundary" in fn:
                # bulk stencils there serve only rank-skin sites (type -2), which a
                # single periodic domain does not have
                assert "if(siteType == -2)" in body
                calls = [c for c in calls if c["kind"] == "boundary"]
            m = re.search(r"lcForce::getBoundaryCase\((.*?)\);", body, re.S)
            case = [t[x] for x in _args(m.group(1))] if m else None
            self.calls[fn] = (calls, case)
        # anchoring: computeBoundaryForcesCPU adds computeBoundaryForce(q, Q of
        # an object neighbour, that object's boundaryObject) for each neighbour
        body = _body(other, "void landauDeGennesLC::computeBoundaryForcesCPU(")
        qn = {name: int(s) for name, s in re.findall(r"\b(\w+) = Qtensors\.data\[neighbors\[(\d)\]\];", body)}
        pairs = re.findall(r"if\(latticeTypes\.data\[neighbors\[(\d)\]\] > 0\)\s*\{\s*computeBoundaryForce\(qCurrent, "
                           r"(\w+), bounds\.data\[latticeTypes\.data\[neighbors\[(\d)\]\]-1\],tempForce\);\s*"
                           r"h_f\.data\[currentIndex\] \+= tempForce;", body)
        assert len(pairs) == 6
        self.anchor_order = []
        for s1, name, s2 in pairs:
            assert int(s1) == int(s2) == qn[name]
            self.anchor_order.append(int(s1))
        # uniform field force and its call sites
        body = _body(other, "void landauDeGennesLC::computeEorHFieldForcesCPU(")
        self.field_exprs = [compile(e, "fieldForce", "eval") for e in
                            [dict(re.findall(r"fieldForce\[(\d)\] = (.*?);", body))[str(c)] for c in range(5)]]
        assert "scalar fieldProduct = anisotropicSusceptibility*vacuumPermeability;" in body
        assert "h_f.data[currentIndex] -= fieldForce;" in body
        assert "if(latticeTypes.data[currentIndex] > 0)//skip boundary sites" in body
        cpu = _body(main, "void landauDeGennesLC::computeForceCPU(")
        self.field_calls = {name: _args(a)[2:5] for name, a in
                            ((n, a) for a in re.findall(r"computeEorHFieldForcesCPU\((.*?)\);", cpu)
                             for n in [_args(a)[2]])}
        assert set(self.field_calls) == {"Efield", "Hfield"}
        # field energy per site (computeEnergyCPU)
        energy = _body(main, "void landauDeGennesLC::computeEnergyCPU(")
        self.e_energy = compile(re.search(r"scalar eFieldAtSite = (.*?);", energy, re.S).group(1), "eE", "eval")
        self.h_energy = compile(re.search(r"scalar hFieldAtSite=(.*?);", energy, re.S).group(1), "eH", "eval")
        # metric
        body = _body(main, "void landauDeGennesLC::correctForceFromMetric(")
        m0 = re.search(r"h_force\.data\[i\]\[0\] = (.*?);", body).group(1)
        m3 = re.search(r"h_force\.data\[i\]\[3\] = (.*?);", body).group(1)
        self.metric_exprs = (compile(m0, "m0", "eval"), compile(m3, "m3", "eval"))

    # ---- the pieces
    def metric(self, f):
        env = {"QxxOld": f[..., 0], "QyyOld": f[..., 3], "twoThirds": 2. / 3.}
        out = f.copy()
        out[..., 0] = eval(self.metric_exprs[0], {}, env)
        out[..., 3] = eval(self.metric_exprs[1], {}, env)
        return out

    def phase(self, q):
        d = np.zeros(15)
        tr = np.zeros(3)
        self.lib.qmref_phase(po.dptr(np.ascontiguousarray(q)), po.dptr(d), po.dptr(tr))
        a, b, c = 0.5 * A, B / 3.0, 0.25 * C
        return -a * d[0:5] - b * d[5:10] - c * d[10:15]

    def spatial(self, fn, L, q, Q, D, types, nbr):
        calls, case = self.calls[fn]
        total = np.zeros(5)
        out = np.zeros(5)
        bcase = None
        if case is not None:
            t6 = np.array([types[nbr[s]] for s in case], np.int32)
            bcase = self.lib.qmref_boundary_case(po.iptr(t6))
        qc = np.ascontiguousarray(q)
        for c in calls:
            k = c["k"]
            if L[k] == 0:
                continue
            nb = np.ascontiguousarray(np.concatenate([Q[nbr[s]] for s in c["nb"]]))
            out[:] = 0
            if k == 1 and c["kind"] == "bulk":
                self.lib.qmref_bulk_l1(L[1], po.dptr(qc), po.dptr(nb), po.dptr(out))
            elif k == 1:
                t6 = np.array([types[nbr[s]] for s in c["types"]], np.int32)
                self.lib.qmref_boundary_l1(L[1], po.dptr(qc), po.dptr(nb), po.iptr(t6), po.dptr(out))
            else:
                nd = np.ascontiguousarray(np.concatenate([D[nbr[s]] for s in c["nd"]]))
                self.lib.qmref_lk(k, L[k], -1 if c["kind"] == "bulk" else bcase, po.dptr(qc), po.dptr(nb),
                                  po.dptr(nd), po.dptr(out))
            total += out
        return total

    def field_force(self, F, aniso, perm):
        env = {"field": types.SimpleNamespace(x=F[0], y=F[1], z=F[2]), "fieldProduct": aniso * perm}
        return np.array([eval(e, {}, env) for e in self.field_exprs])


def _scene(zero_w=False):
    o = po.OracleSystem(*SHAPE)
    o.random_init(77, S0)
    for kind, where, btype, W in OBJECTS:
        W = 0.0 if zero_w else W
        if kind == "wall":
            o.wall(2, where, btype, W, S0)
        else:
            o.sphere(where[0], where[1], btype, W, S0)
    return o


@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("multi", [False, True])
def test_assembled_forces_match_reference_pieces(semantics, multi):
    """Every site's force from the oracle equals the reference's compiled
    per-site functions assembled with the reference's own wiring, for the
    one-constant and the L1-L6 paths, both metric semantics, walls and spheres
    of both anchoring kinds, and E and H fields."""
    ref = Reference()
    L = {1: 2.32, 2: 1.0, 3: 0.5, 4: 0.2, 6: 0.3} if multi else {1: 2.32, 2: 0, 3: 0, 4: 0, 6: 0}
    o = _scene()
    o.set_params(A, B, C, L[1], L[2], L[3], L[4], L[6], multi=multi, semantics=semantics)
    o.set_efield(E["field"], E["eps"], E["eps0"], E["deps"])
    o.set_hfield(H["field"], H["chi"], H["mu0"], H["dchi"])
    got = o.forces()
    Q = o.q()
    D = o.first_derivatives()
    types_ = o.types()
    nbrs = o.neighbors().reshape(-1, 6)
    # Efield, deltaEpsilon, epsilon0 / Hfield, deltaChi, mu0 as the call sites pass them
    members = {"Efield": E["field"], "deltaEpsilon": E["deps"], "epsilon0": E["eps0"],
               "Hfield": H["field"], "deltaChi": H["dchi"], "mu0": H["mu0"]}
    f_field = sum(ref.field_force(members[c[0]], members[c[1]], members[c[2]]) for c in ref.field_calls.values())
    bulk_fn = "computeAllDistortionTermsBulkCPU" if multi else "computeL1BulkCPU"
    bnd_fn = "computeAllDistortionTermsBoundaryCPU" if multi else "computeL1BoundaryCPU"
    want = np.zeros_like(got)
    for i in range(o.n):
        t = types_[i]
        if t > 0:
            continue
        q = Q[i]
        raw = ref.phase(q) - ref.spatial(bulk_fn if t == 0 else bnd_fn, L, q, Q, D, types_, nbrs[i])
        anch = np.zeros(5)
        if t < 0:
            for s in ref.anchor_order:
                tn = types_[nbrs[i][s]]
                if tn > 0:
                    _, _, btype, W = OBJECTS[tn - 1]
                    pf = np.zeros(5)
                    ref.lib.qmref_anchoring_force(btype, W, S0, po.dptr(np.ascontiguousarray(q)),
                                                  po.dptr(np.ascontiguousarray(Q[nbrs[i][s]])), po.dptr(pf))
                    anch += pf
        if semantics == 0 and t == 0:
            # CPU path: the type-0 pass is metric-corrected, then the type-1 pass
            # adds the fields and corrects again (multirankSimulation.cpp:462-466)
            want[i] = ref.metric(ref.metric(raw) - f_field)
        else:
            want[i] = ref.metric(raw + anch - f_field)
    scale = np.abs(want).max()
    err = np.abs(got - want).max()
    assert err <= 1e-13 * scale, "max |dF| %g of %g" % (err, scale)
    assert np.count_nonzero(types_ < 0) > 50 and len(set(types_[types_ > 0])) == len(OBJECTS)


def _fd_gradient(fn, q, h):
    g = np.zeros(5)
    for c in range(5):
        qp, qm = q.copy(), q.copy()
        qp[c] += h
        qm[c] -= h
        g[c] = (fn(qp) - fn(qm)) / (2 * h)
    return g


def test_anchoring_and_field_forces_are_minus_their_energy_gradients():
    """Central differences of the oracle's anchoring and field energies at a
    site equal minus its anchoring and field forces (taken as the difference of
    the oracle's forces with and without the term, metric removed).  The field
    energy is also the reference's expression read from its source; both terms
    are local, so one site's components carry the whole gradient."""
    ref = Reference()
    o = _scene()
    o.set_params(A, B, C, 2.32, semantics=1)
    base = o.forces()
    Q = o.q()
    types_ = o.types()
    minv = np.linalg.inv(np.array([[4 / 3, 0, 0, -2 / 3, 0], [0, 1, 0, 0, 0], [0, 0, 1, 0, 0],
                                   [-2 / 3, 0, 0, 4 / 3, 0], [0, 0, 0, 0, 1]]))
    # metric as a matrix, checked against the reference's expressions
    probe = np.random.default_rng(1).standard_normal((4, 5))
    np.testing.assert_allclose(ref.metric(probe), probe @ np.linalg.inv(minv).T, rtol=1e-15, atol=1e-15)
    rng = np.random.default_rng(8)
    # anchoring: the same scene with every W set to zero removes only the pair forces
    o0 = _scene(zero_w=True)
    assert np.array_equal(o0.types(), types_) and np.array_equal(o0.q(), Q)
    o0.set_params(A, B, C, 2.32, semantics=1)
    f_anch_total = (base - o0.forces()) @ minv.T   # before the metric
    sites = rng.choice(np.flatnonzero(types_ < 0), 12, replace=False)

    def anch_energy(i, qi):
        qq = Q.copy()
        qq[i] = qi
        return o.energy(qq)[1][2]
    for i in sites:
        g = _fd_gradient(lambda qi: anch_energy(i, qi), Q[i].copy(), 1e-5)
        np.testing.assert_allclose(-g, f_anch_total[i], rtol=1e-7, atol=1e-7 * np.abs(f_anch_total[i]).max())
    assert np.abs(f_anch_total[types_ == 0]).max() == 0.0
    # fields
    for which, prm in (("E", E), ("H", H)):
        of = _scene()
        of.set_params(A, B, C, 2.32, semantics=1)
        if which == "E":
            of.set_efield(prm["field"], prm["eps"], prm["eps0"], prm["deps"])
        else:
            of.set_hfield(prm["field"], prm["chi"], prm["mu0"], prm["dchi"])
        f_fld = (of.forces() - base) @ minv.T
        fx, fy, fz = prm["field"]
        if which == "E":
            env = {"Efield": types.SimpleNamespace(x=fx, y=fy, z=fz), "epsilon": prm["eps"],
                   "epsilon0": prm["eps0"], "deltaEpsilon": prm["deps"]}
            expr, comp = ref.e_energy, 3
        else:
            env = {"Hfield": types.SimpleNamespace(x=fx, y=fy, z=fz), "Chi": prm["chi"], "mu0": prm["mu0"],
                   "deltaChi": prm["dchi"]}
            expr, comp = ref.h_energy, 4

        def site_energy(qi):
            return eval(expr, {}, dict(env, qCurrent=qi))
        lc = types_ <= 0
        e_ref = sum(site_energy(Q[i]) for i in np.flatnonzero(lc))
        np.testing.assert_allclose(of.energy()[1][comp], e_ref, rtol=1e-13)
        for i in rng.choice(np.flatnonzero(lc), 6, replace=False):
            g = _fd_gradient(site_energy, Q[i].copy(), 1e-3)   # linear in Q: exact up to rounding
            np.testing.assert_allclose(-g, f_fld[i], rtol=1e-9, atol=1e-12)
        assert np.abs(f_fld[types_ > 0]).max() == 0.0


@pytest.mark.parametrize("wall_type", [0, 1])
def test_object_force_on_a_flat_wall_closed_form(wall_type):
    """computeObjectForces (landauDeGennesLCOtherForces.cpp:147-182) in closed
    form: between walls at z = 0 and z = 7 the liquid crystal holds
    Q(z) = Q0 + (z - 1) G, uniform in x and y, with G diagonal (Gxx, Gyy).  Each of
    the Lx Ly surface sites of the lower wall (z = 1) then has no x or y
    derivative, the one-sided z derivative G, and the area vector (0, 0, -1), so
    the force on the wall is (0, 0, -Lx Ly sigma_zz) with the Ericksen stress
    sigma_zz = e - L1 Tr(G^2) and e the site's energy density: phase energy,
    the pair anchoring energy with the wall and (L1/2) Tr(G^2).  The phase and
    anchoring energies are the reference's compiled functions; the upper wall
    gets +Lx Ly sigma_zz at z = 6.  (For a diagonal G the reference's stress
    equals the Ericksen stress; for off-diagonal gradient components its
    weights are twice the Ericksen ones, which the oracle reproduces.)"""
    lib = po.ref_lib()
    lx, ly, lz = 6, 5, 8
    L1 = 2.32
    o = po.OracleSystem(lx, ly, lz)
    o.random_init(5, S0)
    o.wall(2, 0, wall_type, 5.3, S0)
    o.wall(2, lz - 1, 1 - wall_type, 2.7, S0)
    o.set_params(A, B, C, L1, semantics=1)
    q0 = np.zeros(5)
    lib.qmref_q_from_director(po.dptr(np.array([0.6, 0.0, 0.8])), S0, po.dptr(q0))
    G = np.array([0.013, 0.0, 0.0, -0.021, 0.0])
    Q = o.q().reshape(lz, ly, lx, 5)
    for z in range(1, lz - 1):
        Q[z] = q0 + (z - 1) * G
    o.set_q(Q.reshape(-1, 5))
    trG2 = 2 * (G[0] ** 2 + G[3] ** 2 + G[0] * G[3])

    def density(z, wall_z, btype, W):
        q = np.ascontiguousarray(Q[z, 0, 0])
        d, tr = np.zeros(15), np.zeros(3)
        lib.qmref_phase(po.dptr(q), po.dptr(d), po.dptr(tr))
        phase = 0.5 * A * tr[0] + B / 3.0 * tr[1] + 0.25 * C * tr[2]
        anch = lib.qmref_anchoring_energy(btype, W, S0, po.dptr(q), po.dptr(np.ascontiguousarray(Q[wall_z, 0, 0])))
        return phase + anch + 0.5 * L1 * trG2

    sz_low = density(1, 0, wall_type, 5.3) - L1 * trG2
    sz_high = density(lz - 2, lz - 1, 1 - wall_type, 2.7) - L1 * trG2
    np.testing.assert_allclose(o.object_force(0), [0.0, 0.0, -lx * ly * sz_low], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o.object_force(1), [0.0, 0.0, lx * ly * sz_high], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("types_xz", [(0, 1), (1, 0)])
def test_object_force_at_a_corner_closed_form(types_xz):
    """computeObjectForces' area vector (landauDeGennesLCOtherForces.cpp:160-176)
    at sites touching objects along two axes, in closed form.  Object 1 is a
    wall at z = 0, object 2 a wall at x = 0, object 3 a wall at z = Lz - 1 (so
    that no surface site of object 1 wraps round in z); a later wall takes the
    sites it shares with an earlier one.  The liquid crystal holds
    Q(x) = Q0 + (x - 1) G, uniform in y and z, so at object 1's surface sites
    (z = 1) only the x derivative is nonzero, G (central, or one-sided next to
    object 2 on either side), and the stress is diagonal: sigma_xx = e - L1 Tr G^2,
    sigma_yy = sigma_zz = e, e the site's energy density.  Every site contributes
    a_z = -1; the sites at x = 1 and at x = Lx - 1 (next to object 2 across the
    periodic boundary) also a_x = -1 and +1, although object 2 is another
    object (any object sets the area vector).  So object 1's force is
    (Ly (e(Lx-1) - e(1)), 0, -Ly sum_x e(x))."""
    lib = po.ref_lib()
    lx, ly, lz = 7, 4, 6
    L1 = 2.32
    bt1, bt2 = types_xz
    o = po.OracleSystem(lx, ly, lz)
    o.random_init(9, S0)
    o.wall(2, 0, bt1, 5.3, S0)
    o.wall(0, 0, bt2, 2.7, S0)
    o.wall(2, lz - 1, 1, 1.9, S0)
    o.set_params(A, B, C, L1, semantics=1)
    t = o.types().reshape(lz, ly, lx)
    assert np.all(t[0, :, 1:] == 1) and np.all(t[:lz - 1, :, 0] == 2) and np.all(t[lz - 1] == 3)
    q0 = np.zeros(5)
    lib.qmref_q_from_director(po.dptr(np.array([0.0, 0.6, 0.8])), S0, po.dptr(q0))
    G = np.array([0.011, 0.0, 0.0, -0.017, 0.0])
    Q = o.q().reshape(lz, ly, lx, 5)
    qb1 = np.ascontiguousarray(Q[0, 0, 3])
    qb2 = np.ascontiguousarray(Q[2, 0, 0])
    for x in range(1, lx):
        Q[1:lz - 1, :, x] = q0 + (x - 1) * G
    o.set_q(Q.reshape(-1, 5))
    trG2 = 2 * (G[0] ** 2 + G[3] ** 2 + G[0] * G[3])

    def density(x):
        q = np.ascontiguousarray(Q[1, 0, x])
        d, tr = np.zeros(15), np.zeros(3)
        lib.qmref_phase(po.dptr(q), po.dptr(d), po.dptr(tr))
        e = 0.5 * A * tr[0] + B / 3.0 * tr[1] + 0.25 * C * tr[2] + 0.5 * L1 * trG2
        e += lib.qmref_anchoring_energy(bt1, 5.3, S0, po.dptr(q), po.dptr(qb1))
        if x in (1, lx - 1):
            e += lib.qmref_anchoring_energy(bt2, 2.7, S0, po.dptr(q), po.dptr(qb2))
        return e

    e = {x: density(x) for x in range(1, lx)}
    want = [ly * (e[lx - 1] - e[1]), 0.0, -ly * sum(e.values())]
    assert abs(want[0]) > 1e-3 * abs(want[2])
    np.testing.assert_allclose(o.object_force(0), want, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("types_12", [(0, 1), (1, 0)])
def test_object_force_between_two_walls_closed_form(types_12):
    """computeObjectForces at sites with objects on both sides of one axis, in
    closed form.  Walls at z = 0 (object 1) and z = 2 (object 2) on a periodic
    lattice of Lz = 4: every liquid-crystal site (z = 1 and z = 3) has an object
    below and above.  The reference then sets the area vector's z component
    last from the upper neighbour, a_z = +1 (landauDeGennesLCOtherForces.cpp:
    175-178), and takes the one-sided z derivative against the lower neighbour,
    G = Q - Q(below), the upper object taking precedence (:124-130), though the
    lower neighbour is an object too.  With a uniform Q0 in the liquid crystal
    the stress is diagonal with sigma_xx = sigma_yy = e and sigma_zz = e - 2 L1
    (G0^2 + 2 G1^2 + 2 G2^2 + G3^2 + 2 G4^2 + G0 G3) (:232-241), e the site's
    energy density: phase + L1/2 Tr G^2 + both anchoring terms.  Both walls have
    the same surface sites, so each takes Lx Ly (sigma_zz(1) + sigma_zz(3)) along
    +z.  Choosing the lower neighbour's area vector or derivative instead changes
    the result."""
    lib = po.ref_lib()
    lx, ly, lz = 5, 4, 4
    L1 = 2.32
    bt1, bt2 = types_12
    o = po.OracleSystem(lx, ly, lz)
    o.random_init(3, S0)
    o.wall(2, 0, bt1, 5.3, S0)
    o.wall(2, 2, bt2, 2.7, S0)
    o.set_params(A, B, C, L1, semantics=1)
    t = o.types().reshape(lz, ly, lx)
    assert np.all(t[0] == 1) and np.all(t[2] == 2) and np.all(t[1] == -1) and np.all(t[3] == -1)
    q0 = np.zeros(5)
    lib.qmref_q_from_director(po.dptr(np.array([0.48, 0.6, 0.64])), S0, po.dptr(q0))
    Q = o.q().reshape(lz, ly, lx, 5)
    qb = {0: np.ascontiguousarray(Q[0, 0, 0]), 2: np.ascontiguousarray(Q[2, 0, 0])}
    assert not np.allclose(qb[0], qb[2])
    Q[1] = q0
    Q[3] = q0
    o.set_q(Q.reshape(-1, 5))

    def sigma_zz(below):
        G = q0 - qb[below]
        tr_g2 = 2 * (G[0] ** 2 + G[3] ** 2 + G[0] * G[3] + G[1] ** 2 + G[2] ** 2 + G[4] ** 2)
        d, tr = np.zeros(15), np.zeros(3)
        q = np.ascontiguousarray(q0)
        lib.qmref_phase(po.dptr(q), po.dptr(d), po.dptr(tr))
        e = 0.5 * A * tr[0] + B / 3.0 * tr[1] + 0.25 * C * tr[2] + 0.5 * L1 * tr_g2
        e += lib.qmref_anchoring_energy(bt1, 5.3, S0, po.dptr(q), po.dptr(qb[0]))
        e += lib.qmref_anchoring_energy(bt2, 2.7, S0, po.dptr(q), po.dptr(qb[2]))
        return e - 2 * L1 * (G[0] ** 2 + 2 * G[1] ** 2 + 2 * G[2] ** 2 + G[3] ** 2 + 2 * G[4] ** 2 + G[0] * G[3])

    want = [0.0, 0.0, lx * ly * (sigma_zz(0) + sigma_zz(2))]
    assert abs(want[2]) > 1e-6
    np.testing.assert_allclose(o.object_force(0), want, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(o.object_force(1), want, rtol=1e-12, atol=1e-13)
