"""Fixed workloads for hardware-counter and kernel-trace profiling.

  python open-qmin_amd/qmin/profile_run.py fire [--size 256] [--precision fp64] [--steps 10] [--tile T]
         [--workload bench|full]
      FIRE iterations on the 256^3 walls + sphere workload (as bench.py), or on
      the full-elasticity workload, with the library's default tile or tile T
      (qm_lattice_set_tile).
  python open-qmin_amd/qmin/profile_run.py copy [--mib 512] [--steps 10]
      a plain device-to-device copy of known size, used to calibrate the
      FETCH_SIZE / WRITE_SIZE counters (bytes the counter reports per byte moved).
  python open-qmin_amd/qmin/profile_run.py minimizers --size 384 [--precision fp64] [--workload full|bench]
      the full-elasticity workload (qmin.workloads.full_elasticity: L1..L4 and
      L6, a uniform E field, four homeotropic spheres), or the bench's walls +
      sphere workload (L1..L3); times per iteration of FIRE, Adam, Nesterov and
      online L-BFGS (memory 8).
  python open-qmin_amd/qmin/profile_run.py warmstart --size 128 [--steps 5]
      the warm-start network's BF16 matrix-core convolutions on an L^3 lattice
      with five spheres: --steps forward passes (inference), then --steps
      weight- and input-gradient passes with a unit output gradient; prints the
      per-layer FLOP counts (2 x 27 x cin x cout per site).
  python open-qmin_amd/qmin/profile_run.py ensemble [--count 1024] [--steps 256]
      the bench's ensemble workload (qmin.workloads.ensemble_specs, FP32, L1):
      the tile autotuner, then --steps batched FIRE iterations over all lattices
      (no tolerance stop), replayed from hipGraphs of 32 iterations.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, REPO)


def ensemble(qmin, count, steps):
    import time
    from qmin import workloads
    A, B, C, S0 = workloads.phase_constants()
    lats = [workloads.build_lattice(qmin, sp, qmin.QM_FP32, S0) for sp in workloads.ensemble_specs(count)]
    sites = sum(l.n for l in lats)
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=qmin.SEMANTICS_REF_GPU)
    ens = qmin.Ensemble(lats)
    tile, zchunk, ms = ens.autotune(force)
    print("%d lattices, %d sites; autotune: tile %d, z chunk %d, %.3f ms per batched pass"
          % (count, sites, tile, zchunk, ms), flush=True)
    fire = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=0.0, max_iter=steps)
    t0 = time.perf_counter()
    its, _ = ens.fire_minimize(fire, force, 32)
    t1 = time.perf_counter()
    updates = sum(l.n * int(i) for l, i in zip(lats, its))
    print("%d iterations each: %.3f s, %.3e site-updates/s" % (steps, t1 - t0, updates / (t1 - t0)), flush=True)


def minimizers(qmin, L, precision, name, workload="full"):
    import time
    if workload == "full":
        from qmin import workloads
        lat, f = workloads.full_elasticity(qmin, L, precision)
    else:
        from bench import build_workload
        lat, f = build_workload(qmin, L, precision)
    print("%d^3 %s, %d boundary sites" % (L, name, lat.n_boundary()), flush=True)
    runs = (("FIRE", qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=1e-12, max_iter=10 ** 9), 20),
            ("Adam", qmin.Adam(dt=1e-4, fc=1e-12, max_iter=10 ** 9), 10),
            ("Nesterov", qmin.Nesterov(dt=1e-4, mu=0.01, fc=1e-12, max_iter=10 ** 9), 10),
            ("LoLBFGS(m=8)", qmin.LoLBFGS(m=8, dt=1e-4, fc=1e-12, max_iter=10 ** 9), 5))
    for label, m, its in runs:
        if label == "FIRE":
            m.run_iterations(f, lat, 3)
            t0 = time.perf_counter()
            m.run_iterations(f, lat, its)
        else:
            m.set_max_iterations(2)
            m.minimize(f, lat)
            m.set_max_iterations(2 + its)
            t0 = time.perf_counter()
            m.minimize(f, lat)
        t1 = time.perf_counter()
        print("%s: %.3f ms per iteration, %.3e site-updates/s" % (label, (t1 - t0) / its * 1e3, lat.n * its / (t1 - t0)),
              flush=True)


def warmstart(L, steps):
    import time
    import numpy as np
    import torch  # noqa: F401  (before the library: one HIP runtime)
    import qmin
    from qmin import warmstart as ws
    from qmin import workloads
    A, B, C, S0 = workloads.phase_constants()
    rng = np.random.default_rng(5)
    lat = qmin.Lattice(L, L, L)
    lat.random_init(1, S0)
    for _ in range(5):
        lat.sphere(tuple(float(v) for v in rng.uniform(0, L, 3)), float(rng.uniform(L / 16, L / 8)), qmin.HOMEOTROPIC,
                   5.0, S0)
    net = ws.WarmStartNet(seed=0)
    feats = ws.lattice_features(lat, (0.0, 0.0, 1.0))
    dims = lat.shape
    for _ in range(steps):
        acts = net.forward(feats, dims)
    torch.cuda.synchronize()
    gout = torch.full((lat.n * 16,), 0x3f80, dtype=torch.int16, device="cuda")   # BF16 1.0
    for _ in range(steps):
        net.backward(acts, gout, dims)
    torch.cuda.synchronize()
    for li, (cin, cout, dil) in enumerate(ws.LAYERS):
        print("layer %d: %d -> %d channels, dilation %d: %.3e FLOP per forward pass" % (
            li, cin, cout, dil, 2.0 * 27 * cin * cout * lat.n), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["fire", "copy", "minimizers", "warmstart", "ensemble"])
    ap.add_argument("--count", type=int, default=1024)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--mib", type=int, default=512)
    ap.add_argument("--tile", type=int, default=-1, help="tile of the FIRE lattice pass (-1: the library's default)")
    ap.add_argument("--workload", choices=["bench", "full"], default=None,
                    help="the bench workload, or the full-elasticity one (L1-L4, L6, E field, four spheres); "
                         "default: bench for fire, full for minimizers")
    args = ap.parse_args()
    if args.what == "copy":
        import torch
        n = args.mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(args.steps):
            b.copy_(a)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / args.steps
        print("copied %d MiB x %d: %.3f ms per copy, %.1f GB/s read+write"
              % (args.mib, args.steps, ms, 2 * args.mib * (1 << 20) / (ms * 1e-3) / 1e9))
        return
    if args.what == "warmstart":
        warmstart(args.size, args.steps)
        return
    import qmin
    if args.what == "ensemble":
        ensemble(qmin, args.count, args.steps)
        return
    if args.what == "minimizers":
        minimizers(qmin, args.size, qmin.QM_FP64 if args.precision == "fp64" else qmin.QM_FP32, args.precision,
                   args.workload or "full")
        return
    from bench import build_workload, fire_params
    precision = qmin.QM_FP64 if args.precision == "fp64" else qmin.QM_FP32
    if (args.workload or "bench") == "full":
        from qmin import workloads
        lat, force = workloads.full_elasticity(qmin, args.size, precision)
    else:
        lat, force = build_workload(qmin, args.size, precision)
    if args.tile >= 0:
        lat.set_tile(args.tile)
    m = qmin.FIRE(fc=1e-8, max_iter=10 ** 9, **fire_params())
    m.run_iterations(force, lat, args.steps)
    print("ran %d FIRE iterations on %d^3 (%s), forceMax %.4g" % (args.steps, args.size, args.precision,
                                                                  m.state().forceMax))


if __name__ == "__main__":
    main()
