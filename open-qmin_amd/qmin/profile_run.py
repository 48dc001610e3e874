"""Fixed workloads for hardware-counter and kernel-trace profiling.

  python open-qmin_amd/qmin/profile_run.py fire [--size 256] [--precision fp64] [--steps 10]
      FIRE iterations on the 256^3 walls + sphere workload (as bench.py).
  python open-qmin_amd/qmin/profile_run.py copy [--mib 512] [--steps 10]
      a plain device-to-device copy of known size, used to calibrate the
      FETCH_SIZE / WRITE_SIZE counters (bytes the counter reports per byte moved).
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, REPO)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["fire", "copy"])
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--mib", type=int, default=512)
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
    import qmin
    from bench import build_workload, fire_params
    precision = qmin.QM_FP64 if args.precision == "fp64" else qmin.QM_FP32
    lat, force = build_workload(qmin, args.size, precision)
    m = qmin.FIRE(fc=1e-8, max_iter=10 ** 9, **fire_params())
    m.run_iterations(force, lat, args.steps)
    print("ran %d FIRE iterations on %d^3 (%s), forceMax %.4g" % (args.steps, args.size, args.precision,
                                                                  m.state().forceMax))


if __name__ == "__main__":
    main()
