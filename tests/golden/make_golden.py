"""Write tests/golden/reference_vectors.npz from the compiled reference.

Every array here is an output of the reference's own code, compiled from its
sources by oracle/Makefile into oracle/_ref/libqmref.so, on seeded inputs:
per-site stencils (lcForces.h), anchoring (landauDeGennesLCBoundary.h), Q
algebra (qTensorFunctions.h), the RNG stream (noiseSource.cpp), the lattice
model (cubicLattice.cpp, qTensorLatticeModel.cpp) and the minimizer loops
(energyMinimizer*.cpp, velocityVerlet.cpp).  The minimizer traces are the
reference's loops driven with forces from the CPU oracle (the force class
itself needs mpi.h and is not compiled here).

Run here, where the reference checkout exists:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import pyoracle as po  # noqa: E402

OUT = os.path.join(REPO, "tests", "golden", "reference_vectors.npz")
CASES = [xx + 4 * yy + 16 * zz for zz in range(3) for yy in range(3) for xx in range(3)]
SMALL = (6, 5, 7)
SPHERE = ((3.0, 2.5, 3.5), 1.8)


def sphere_sites(shape, center, radius):
    """Lattice indices in createSphericalColloid's loop order (multirankSimulationBoundaries.cpp:169-215)."""
    out = []
    lx, ly, lz = shape
    x = math.ceil(center[0] - radius)
    while x < math.floor(center[0] + radius):
        y = math.ceil(center[1] - radius)
        while y < math.floor(center[1] + radius):
            z = math.ceil(center[2] - radius)
            while z < math.floor(center[2] + radius):
                d = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
                if d < radius * radius:
                    out.append(x % lx + lx * ((y % ly) + ly * (z % lz)))
                z += 1
            y += 1
        x += 1
    return np.array(out, np.int32)


def small_scene(params, semantics, efield=None):
    """Oracle system and reference model with the same walls/sphere on the small lattice."""
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    o = po.OracleSystem(*SMALL)
    o.random_init(1234, S0)
    o.wall(2, 0, 1, 5.3, S0)
    o.sphere(SPHERE[0], SPHERE[1], 0, 5.3, S0)
    o.set_params(A, B, C, *params, semantics=semantics)
    if efield is not None:
        o.set_efield(efield, 1.0, 1.0, 1.0)
    r = po.RefModel(*SMALL)
    r.set_q(o.q())
    lx, ly, _ = SMALL
    wall = np.array([x + lx * (y + ly * 0) for x in range(lx) for y in range(ly)], np.int32)
    r.add_object(wall, 1, 5.3, S0)
    r.add_object(sphere_sites(SMALL, *SPHERE), 0, 5.3, S0)
    return o, r


FIRE_CASES = [
    ("one", (2.32, 0.0, 0.0, 0.0, 0.0), 0, None),
    ("multi", (2.32, 0.5, 0.3, 0.0, 0.0), 0, None),
    ("all", (2.32, 1.0, 0.5, 0.2, 0.3), 0, (0.0, 0.0, 0.5)),
    ("multi_gpu", (2.32, 0.5, 0.3, 0.0, 0.0), 1, None),
]


def main():
    if not po.ref_available():
        sys.exit("build oracle/_ref first: make -C oracle ref")
    ref = po.ref_lib()
    rng = np.random.default_rng(7)
    g = {}
    D = po.dptr

    # per-site stencils
    X = rng.standard_normal((16, 125))
    g["stencil_x"] = X
    for k in (2, 3, 4, 6):
        for c in [-1] + CASES:
            out = np.zeros((16, 5))
            for i in range(16):
                x = np.ascontiguousarray(X[i])
                ref.qmref_lk(k, 1.7, c, D(x[:5]), D(x[5:35]), D(x[35:]), D(out[i]))
            g["stencil_L%d_%s" % (k, "bulk" if c < 0 else c)] = out
    # L1
    T = rng.integers(-1, 3, size=(16, 6)).astype(np.int32)
    g["l1_types"] = T
    l1b, l1e = np.zeros((16, 5)), np.zeros((16, 5))
    for i in range(16):
        x = np.ascontiguousarray(X[i])
        ref.qmref_bulk_l1(2.32, D(x[:5]), D(x[5:35]), D(l1b[i]))
        t = np.ascontiguousarray(T[i])
        ref.qmref_boundary_l1(2.32, D(x[:5]), D(x[5:35]), po.iptr(t), D(l1e[i]))
    g["l1_bulk"], g["l1_boundary"] = l1b, l1e
    # phase
    Q = 0.3 * rng.standard_normal((32, 5))
    g["phase_q"] = Q
    ph, tr = np.zeros((32, 15)), np.zeros((32, 3))
    for i in range(32):
        ref.qmref_phase(D(np.ascontiguousarray(Q[i])), D(ph[i]), D(tr[i]))
    g["phase_grad"], g["phase_tr"] = ph, tr
    # anchoring
    QB = 0.3 * rng.standard_normal((32, 5))
    nu = rng.standard_normal((32, 3))
    nu /= np.linalg.norm(nu, axis=1, keepdims=True)
    QBP = np.zeros((32, 5))
    QBP[:, :3] = nu
    g["anch_q"], g["anch_qb_h"], g["anch_qb_p"] = Q, QB, QBP
    for name, qb, bt in (("h", QB, 0), ("p", QBP, 1)):
        fo, eo = np.zeros((32, 5)), np.zeros(32)
        for i in range(32):
            q, b = np.ascontiguousarray(Q[i]), np.ascontiguousarray(qb[i])
            ref.qmref_anchoring_force(bt, 5.3, 0.53, D(q), D(b), D(fo[i]))
            eo[i] = ref.qmref_anchoring_energy(bt, 5.3, 0.53, D(q), D(b))
        g["anch_force_" + name], g["anch_energy_" + name] = fo, eo
    # eigenvalues
    ev = np.zeros((32, 3))
    for i in range(32):
        ref.qmref_eigenvalues(D(np.ascontiguousarray(Q[i])), D(ev[i]))
    g["eig_q"], g["eig_ev"] = Q, ev
    # RNG stream
    u = np.zeros(64)
    ref.qmref_uniform_stream(1234, 64, D(u))
    g["uniform_1234"] = u
    # model: neighbours, init, types, defects
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    r = po.RefModel(*SMALL)
    g["small_shape"] = np.array(SMALL)
    g["small_neighbors"] = r.neighbors()
    r.random_init(1234, S0)
    g["small_init_q"] = r.q()
    lx, ly, _ = SMALL
    wall = np.array([x + lx * (y + ly * 0) for x in range(lx) for y in range(ly)], np.int32)
    r.add_object(wall, 1, 5.3, S0)
    r.add_object(sphere_sites(SMALL, *SPHERE), 0, 5.3, S0)
    g["small_types"] = r.types()
    g["small_defects"] = r.defects(0)
    # minimizers (reference loops, oracle forces)
    for name, params, sem, ef in FIRE_CASES:
        o, r = small_scene(params, sem, ef)
        fn, ctx = o.force_callback()
        it, tr = r.fire(fn, ctx, max_iter=40)
        g["fire_%s_trace" % name] = tr
        g["fire_%s_q" % name] = r.q()
    o, r = small_scene((2.32, 0.5, 0.3, 0.0, 0.0), 0)
    fn, ctx = o.force_callback()
    g["adam_trace"] = r.adam(fn, ctx, dt=1e-3, max_iter=30)[1]
    g["adam_q"] = r.q()
    o, r = small_scene((2.32, 0.5, 0.3, 0.0, 0.0), 0)
    fn, ctx = o.force_callback()
    g["nesterov_trace"] = r.nesterov(fn, ctx, dt=1e-2, mu=0.5, scheduled=True, max_iter=30)[1]
    g["nesterov_q"] = r.q()
    o, r = small_scene((2.32, 0.5, 0.3, 0.0, 0.0), 0)
    fn, ctx = o.force_callback()
    g["lolbfgs_trace"] = r.lolbfgs(fn, ctx, m=8, dt=1e-2, max_iter=30)[1]
    g["lolbfgs_q"] = r.q()
    np.savez_compressed(OUT, **g)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
