"""Train the warm-start network on converged lattices from the ensemble
minimizer and measure what it buys on held-out geometries.

  python open-qmin_amd/qmin/warmstart_experiment.py [--train 4096] [--epochs 8]
         [--weights out.npz] [--load weights.npz] [--skip-256]

Training set: `--train` 64^3 lattices (half with z walls of either anchoring,
0-3 spheres, a seeded uniform initial director; L1 = 2.32, L2 = 0.5, L3 = 0.3
as the evaluation) minimized together in FP32 by the
ensemble minimizer to forceMax < 1e-7 (at most 4000 FIRE iterations), with the
converged Q and the far-field director (principal axis of the mean bulk Q) as
the target and the conditioning input.

Evaluation, each in FP64 to forceMax < 1e-8 with FIRE:
  * the 256^3 walls + sphere workload of bench.py (far-field director x);
  * a held-out 128^3 periodic lattice with 5 spheres (far-field director z);
from (a) the seeded random start, (b) a uniform director along the far-field
director, (c) the network's prediction; time includes the prediction.
Prints one JSON line.
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


def eval_geometry(name, build, director, fire_kw, net, fc=1e-8):
    A, B, C, S0 = workloads.phase_constants()
    force = build.force
    out = {}
    for mode in ("random", "uniform", "network"):
        lat = build(qmin.QM_FP64, director if mode == "uniform" else None)
        t0 = time.perf_counter()
        if mode == "network":
            net.predict(lat, director)
        m = qmin.FIRE(fc=fc, max_iter=400000, **fire_kw)
        st = m.minimize(force, lat)
        secs = time.perf_counter() - t0
        e, _ = force.compute_energy(lat)
        out[mode] = {"iterations": st.iterations, "seconds": secs, "forceMax": st.forceMax,
                     "converged": bool(st.converged), "energy_per_site": e / lat.n_active()}
        log("%s %s: %d iterations, %.2f s, E/N %.12f" % (name, mode, st.iterations, secs, e / lat.n_active()))
        del lat
    e0 = out["random"]["energy_per_site"]
    for mode in ("uniform", "network"):
        out[mode]["speedup_vs_random"] = out["random"]["seconds"] / out[mode]["seconds"]
        out[mode]["energy_rel_diff_vs_random"] = (out[mode]["energy_per_site"] - e0) / abs(e0)
    return out


class Geometry:
    def __init__(self, dims, seed, walls, spheres, force):
        self.dims, self.seed, self.walls, self.spheres, self.force = dims, seed, walls, spheres, force

    def add_objects(self, lat):
        A, B, C, S0 = workloads.phase_constants()
        for xyz, plane, wt, W in self.walls:
            lat.wall(xyz, plane, wt, W, S0)
        for c, r, t, W in self.spheres:
            lat.sphere(c, r, t, W, S0)

    def __call__(self, precision, uniform=None):
        A, B, C, S0 = workloads.phase_constants()
        lat = qmin.Lattice(*self.dims, precision=precision)
        if uniform is None:
            lat.random_init(self.seed, S0)
        else:
            lat.uniform_init(uniform, S0)
        self.add_objects(lat)
        return lat


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--weights", default="")
    ap.add_argument("--load", default="")
    ap.add_argument("--skip-256", action="store_true")
    args = ap.parse_args()
    A, B, C, S0 = workloads.phase_constants()
    result = {}
    net = ws.WarmStartNet(seed=0)
    if args.load:
        net.load(args.load)
    else:
        # the same physics as the evaluation geometries
        force32 = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
        t0 = time.perf_counter()
        samples = ws.make_dataset(qmin, ws.training_specs(args.train, 64), force32, log=log)
        t1 = time.perf_counter()
        hist = net.train(samples, epochs=args.epochs, lr=args.lr, log=log)
        t2 = time.perf_counter()
        result["training"] = {"lattices": args.train, "size": 64, "data_seconds": t1 - t0, "train_seconds": t2 - t1,
                              "epochs": args.epochs, "mse_per_epoch": hist}
        if args.weights:
            net.save(args.weights)
        del samples
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    rng = np.random.default_rng(99)
    spheres = []
    for _ in range(5):
        c = tuple(float(v) for v in rng.uniform(0, 128, 3))
        spheres.append((c, float(rng.uniform(8, 16)), int(rng.integers(0, 2)), float(rng.uniform(1, 10))))
    g128 = Geometry((128, 128, 128), 4321, [], spheres, force)
    result["eval_128_5_spheres"] = eval_geometry("128^3", g128, (0.0, 0.0, 1.0), dict(dt=5e-4, dt_max=2e-2), net)
    if not args.skip_256:
        g256 = Geometry((256, 256, 256), 1234, [(2, 0, 1, 5.3), (2, 255, 1, 5.3)],
                        [((128.0, 128.0, 128.0), 20.0, 0, 5.3)], force)
        result["eval_256_walls_sphere"] = eval_geometry("256^3", g256, (1.0, 0.0, 0.0), dict(dt=5e-4, dt_max=5e-2),
                                                        net)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
