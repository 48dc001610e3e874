"""Train the warm-start network on converged lattices from the ensemble
minimizer and measure what it buys on held-out geometries.

  python open-qmin_amd/qmin/warmstart_experiment.py [--train 2048] [--train-walls 0] [--train-diagonal 0]
         [--epochs 6] [--lr 5e-4] [--energy-weight 0.1] [--max-iter 6000]
         [--init weights.npz] [--weights out.npz] [--load weights.npz] [--skip-256]

Training set: `--train` 64^3 periodic lattices as the ensemble workload's (0-3
spheres of radius 4-16, homeotropic or degenerate planar, W log-uniform in
[0.5, 20]) from seeded random-director starts (L1 = 2.32, L2 = 0.5, L3 = 0.3 as
the evaluation), `--train-walls` 64^3 lattices of warmstart.training_specs
(half with z walls, uniform starts) and `--train-diagonal` lattices of
warmstart.diagonal_start_specs (the ensemble's geometries, uniform starts along
a body diagonal), minimized together in FP32 by the ensemble
minimizer to forceMax < 1e-7 (at most `--max-iter` FIRE iterations); the
converged Q is the target, the principal axis of its mean bulk Q the far-field
input.  Loss: mean squared error to the target plus `--energy-weight` times the
Landau-de Gennes energy of the prediction per liquid-crystal component (its
gradient is the library's force on the predicted state).  `--init` starts the
training from saved weights; `--load` skips it.

Evaluation, FP32 state with FP64 reductions, FIRE to forceMax < 1e-8 from (a)
seeded random starts, (b) a uniform director along a body diagonal, (c) the
network's prediction along the same diagonal (time includes the prediction):
held-out periodic geometries with 2-6 spheres (warmstart.heldout_geometry) at
128^3 (three seeds) and 256^3 (one seed).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402
import torch  # noqa: E402,F401  (before the library: one HIP runtime)

import qmin  # noqa: E402
from qmin import warmstart as ws  # noqa: E402
from qmin import workloads  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build(L, spheres, precision, start, seed=4321, director=None):
    A, B, C, S0 = workloads.phase_constants()
    lat = qmin.Lattice(L, L, L, precision=precision)
    if start == "random":
        lat.random_init(seed, S0)
    else:
        lat.uniform_init(director, S0)
    for c, r, t, W in spheres:
        lat.sphere(c, r, t, W, S0)
    return lat


def eval_geometry(name, L, spheres, net, force, random_seeds=(4321, 99), fc=1e-8, precision=qmin.QM_FP32):
    d = ws.diagonal_director()
    A, B, C, S0 = workloads.phase_constants()
    out = {"spheres": len(spheres)}
    runs = [("random_%d" % sd, "random", sd) for sd in random_seeds] + [("uniform", "uniform", 0),
                                                                        ("network", "network", 0)]
    for tag, start, sd in runs:
        lat = build(L, spheres, precision, "random" if start == "random" else "uniform", sd, d)
        t0 = time.perf_counter()
        if start == "network":
            net.predict(lat, d, S0)
        st = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=fc, max_iter=400000).minimize(force, lat)
        secs = time.perf_counter() - t0
        e, _ = force.compute_energy(lat)
        out[tag] = {"iterations": st.iterations, "seconds": secs, "forceMax": st.forceMax,
                    "converged": bool(st.converged), "energy_per_site": e / lat.n_active()}
        log("%s %s: %d iterations, %.2f s, E/N %.12f" % (name, tag, st.iterations, secs, e / lat.n_active()))
        del lat
    rnd = [out["random_%d" % sd] for sd in random_seeds]
    t_rnd = float(np.mean([r["seconds"] for r in rnd]))
    for tag in ("uniform", "network"):
        out[tag]["speedup_vs_random_mean"] = t_rnd / out[tag]["seconds"]
        out[tag]["energy_rel_diff_vs_random"] = [(out[tag]["energy_per_site"] - r["energy_per_site"]) /
                                                 abs(r["energy_per_site"]) for r in rnd]
    out["network"]["speedup_vs_uniform"] = out["uniform"]["seconds"] / out["network"]["seconds"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2048)
    ap.add_argument("--train-walls", type=int, default=0)
    ap.add_argument("--train-diagonal", type=int, default=0)
    ap.add_argument("--init", default="")
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--energy-weight", type=float, default=0.1)
    ap.add_argument("--max-iter", type=int, default=6000)
    ap.add_argument("--weights", default="")
    ap.add_argument("--load", default="")
    ap.add_argument("--skip-256", action="store_true")
    args = ap.parse_args()
    A, B, C, S0 = workloads.phase_constants()
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    result = {}
    net = ws.WarmStartNet(seed=0)
    if args.load:
        net.load(args.load)
    else:
        if args.init:
            net.load(args.init)
        t0 = time.perf_counter()
        specs = ws.random_start_specs(args.train, 64)
        if args.train_walls:
            specs += ws.training_specs(args.train_walls, 64)
        if args.train_diagonal:
            specs += ws.diagonal_start_specs(args.train_diagonal, 64)
        samples = ws.make_dataset(qmin, specs, force, max_iter=args.max_iter, log=log,
                                  keep_lattices=args.energy_weight > 0)
        t1 = time.perf_counter()
        hist = net.train(samples, epochs=args.epochs, lr=args.lr, log=log, force=force,
                         energy_weight=args.energy_weight)
        t2 = time.perf_counter()
        result["training"] = {"lattices": args.train, "wall_lattices": args.train_walls,
                              "diagonal_lattices": args.train_diagonal, "size": 64,
                              "init": os.path.basename(args.init), "data_seconds": t1 - t0,
                              "train_seconds": t2 - t1, "epochs": args.epochs, "energy_weight": args.energy_weight,
                              "mse_per_epoch": hist}
        if args.weights:
            net.save(args.weights)
        del samples
    for seed in (11, 12, 13):
        result["heldout_128_seed%d" % seed] = eval_geometry("128^3 seed %d" % seed, 128,
                                                            ws.heldout_geometry(128, seed), net, force)
    if not args.skip_256:
        result["heldout_256_seed21"] = eval_geometry("256^3 seed 21", 256, ws.heldout_geometry(256, 21), net, force,
                                                     random_seeds=(4321,))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
