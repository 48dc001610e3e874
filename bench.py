"""FIRE throughput on one MI355X: lattice site-updates per second.

Workload (the reference's headline single-GPU configuration): a 256^3 lattice,
periodic in x and y, degenerate-planar walls (W = 5.3) at z = 0 and z = 255, a
homeotropic sphere (W = 5.3, radius 20) at the centre, bulk constants
A = -0.172, B = -2.12, C = 1.73 (rescaled as the openQmin driver does),
L1 = 2.32, L2 = 0.5, L3 = 0.3, seeded random directors (seed 1234).  One step
is one FIRE iteration (velocity Verlet + force + FIRE update) over all 256^3
sites; value = sites x steps / wall time of the timed steps, with the state
resident in HBM.

Also reported: the fused lattice kernel's average duration from HIP events and
its achieved algorithmic bandwidth against the 8 TB/s HBM peak, and the CPU
oracle (a restatement of the reference CPU path) timed on one host core on a
bounded sample of the same physics.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "open-qmin_amd"))

METRIC = "FIRE site-updates/s + time-to-max|F|<1e-8 at 256³, FP64 vs mixed vs warm-start"
PEAK_HBM_GBS = 8000.0
BYTES_PER_SITE = {0: 241.0, 1: 121.0}  # read+write Q, v, F (5 components each) + 1 B type


def phase_constants():
    a, b, c = -0.172, -2.12, 1.73
    A, B, C = -1.0, -b / a, -c / a
    S0 = (-B + (B * B - 24 * A * C) ** 0.5) / (6 * C)
    return A, B, C, S0


def build_workload(qmin, L, precision, seed=1234, uniform=None):
    A, B, C, S0 = phase_constants()
    lat = qmin.Lattice(L, L, L, precision=precision)
    if uniform is None:
        lat.random_init(seed, S0)
    else:
        lat.uniform_init(uniform, S0)
    lat.wall(2, 0, qmin.DEGENERATE_PLANAR, 5.3, S0)
    lat.wall(2, L - 1, qmin.DEGENERATE_PLANAR, 5.3, S0)
    c = L / 2.0
    lat.sphere((c, c, c), 20.0 * L / 256.0, qmin.HOMEOTROPIC, 5.3, S0)
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    return lat, force


def fire_params():
    dt = 5e-4
    return dict(dt=dt, alpha_start=0.99, dt_max=100 * dt, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9, n_min=4,
                alpha_min=0.0)


def gpu_throughput(qmin, L, precision, steps, warmup, zchunk=0):
    lat, force = build_workload(qmin, L, precision)
    lat.set_zchunk(zchunk)
    m = qmin.FIRE(fc=1e-8, max_iter=10 ** 9, **fire_params())
    if warmup:
        m.run_iterations(force, lat, warmup)
    t0 = time.perf_counter()
    m.run_iterations(force, lat, steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    # kernel durations with HIP events on the library's stream (separate run of the same steps)
    qmin.timing_reset()
    qmin.timing_enable(True)
    m.run_iterations(force, lat, min(steps, 20))
    qmin.timing_enable(False)
    n0, ms0 = qmin.timing_get(0)
    n1, ms1 = qmin.timing_get(1)
    st = m.state()
    return dict(wall=wall, n=lat.n, march_ms=ms0 / max(n0, 1), boundary_ms=ms1 / max(n1, 1),
                forceMax=st.forceMax, iterations=st.iterations, n_boundary=lat.n_boundary())


# FP32 rounding floor of forceMax on the 256^3 workload: 2.2e-10 (a 30,000-
# iteration FP32 trace levels off there).  The mixed runs aiming below it leave
# FP32 at 3e-10, where the FP32 trajectory still descends at full speed.
MIXED_SWITCH = 3e-10


def tolerance_chain(qmin, L, mode, fcs, switch_fc=None, max_iter=400000, ring=False):
    """FIRE from the seeded random start, stopping at each forceMax tolerance of
    fcs in turn (the reference's definition, energyMinimizerFIRE.cpp:110): one
    run, the minimizer continuing with its dt, alpha and N+ after each stop, so
    each stage's time is what a run to that tolerance takes.  mode "fp64": all
    FP64; "mixed": FP32 state and stencils with FP64 reductions until forceMax <
    max(fcs[0], switch_fc), then the state is converted on the device and the
    same minimizer polishes in FP64.  The energy (and with ring=True the defect
    ring of the first stage) is taken outside the timed region."""
    prec = qmin.QM_FP64 if mode == "fp64" else qmin.QM_FP32
    lat, force = build_workload(qmin, L, prec)
    first_fc = fcs[0] if mode == "fp64" else max(fcs[0], switch_fc or fcs[0])
    m = qmin.FIRE(fc=first_fc, max_iter=max_iter, **fire_params())
    elapsed = 0.0
    t0 = time.perf_counter()
    st = m.minimize(force, lat)
    first = st.iterations
    fp32_seconds = time.perf_counter() - t0
    elapsed += fp32_seconds
    if mode == "mixed":
        t0 = time.perf_counter()
        lat.set_precision(qmin.QM_FP64)
        elapsed += time.perf_counter() - t0
    out = []
    for k, fc in enumerate(fcs):
        if mode == "mixed" or k > 0:
            m.set_force_cutoff(fc)
            m.set_max_iterations(m.state().iterations + max_iter)
            t0 = time.perf_counter()
            st = m.minimize(force, lat)
            elapsed += time.perf_counter() - t0
        e, _ = force.compute_energy(lat)
        r = dict(forceMax_below=fc, seconds=elapsed, iterations=st.iterations, forceMax=st.forceMax,
                 converged=bool(st.converged), energy_per_site=e / lat.n_active())
        if mode == "mixed":
            r.update(fp32_until=first_fc, fp32_iterations=first, fp32_seconds=fp32_seconds)
        if ring and k == 0:
            from qmin import defects
            c = L / 2.0
            r["defect_ring"] = defects.ring_locus(lat, (c, c, c), 20.0 * L / 256.0)
            # the reference's forceMax is |F|_2 / N; the largest force on any one site, for comparison
            import numpy as np
            r["max_site_force"] = float(np.sqrt((lat.f() ** 2).sum(axis=1)).max())
        print("time to forceMax<%g (%s): %d iterations, %.2f s" % (fc, mode, st.iterations, elapsed),
              file=sys.stderr, flush=True)
        out.append(r)
    return out


def time_to_tolerance(qmin, L, mode, fc=1e-8, switch_fc=None, max_iter=400000):
    """One run of tolerance_chain to a single tolerance."""
    return tolerance_chain(qmin, L, mode, (fc,), switch_fc, max_iter)[0]


WEIGHTS = os.path.join(REPO, "open-qmin_amd", "warmstart", "weights_64.npz")


def warm_start(qmin, L, fc=1e-8, weights=WEIGHTS):
    """FIRE to forceMax < fc on the bench workload from the trained network's
    prediction (committed weights) and from a uniform director along the same
    far-field direction (x, in the plane of the planar walls), in FP64; and from
    the prediction in mixed precision (FP32 to fc, then the FP64 polish as in
    time_to_tolerance).  Times include the prediction."""
    from qmin import warmstart as ws
    net = ws.WarmStartNet(seed=0)
    net.load(weights)
    A, B, C, S0 = phase_constants()
    out = {}
    for mode in ("network", "uniform_director", "network_mixed"):
        mixed = mode == "network_mixed"
        lat, force = build_workload(qmin, L, qmin.QM_FP32 if mixed else qmin.QM_FP64,
                                    uniform=None if mode.startswith("network") else (1.0, 0.0, 0.0))
        m = qmin.FIRE(fc=fc, max_iter=400000, **fire_params())
        t0 = time.perf_counter()
        if mode.startswith("network"):
            net.predict(lat, (1.0, 0.0, 0.0), S0)
        st = m.minimize(force, lat)
        first = st.iterations
        if mixed:
            lat.set_precision(qmin.QM_FP64)
            m.set_max_iterations(first + 400000)
            st = m.minimize(force, lat)
        t1 = time.perf_counter()
        e, _ = force.compute_energy(lat)
        out[mode] = dict(seconds=t1 - t0, iterations=st.iterations, forceMax=st.forceMax,
                         converged=bool(st.converged), energy_per_site=e / lat.n_active())
        if mixed:
            out[mode].update(fp32_iterations=first, polish_iterations=st.iterations - first)
        print("time to forceMax<%g (%s start): %d iterations, %.2f s" % (fc, mode, st.iterations, t1 - t0),
              file=sys.stderr, flush=True)
    return out


def warm_start_energy_matched(qmin, L, target, chunk=25, max_iter=20000, weights=WEIGHTS):
    """Time from the network's prediction, and from the uniform director, until
    the energy per site reaches target (the random start's final energy at the
    force tolerance): FP64 FIRE on the bench workload in runs of chunk
    iterations, the energy taken between runs outside the timed region; the time
    includes the prediction.  The minimizer does not stop on forceMax here."""
    from qmin import warmstart as ws
    net = ws.WarmStartNet(seed=0)
    net.load(weights)
    A, B, C, S0 = phase_constants()
    out = {"target_energy_per_site": target, "chunk": chunk}
    for mode in ("network", "uniform_director"):
        lat, force = build_workload(qmin, L, qmin.QM_FP64,
                                    uniform=None if mode == "network" else (1.0, 0.0, 0.0))
        m = qmin.FIRE(fc=0.0, max_iter=10 ** 9, **fire_params())
        elapsed = 0.0
        t0 = time.perf_counter()
        if mode == "network":
            net.predict(lat, (1.0, 0.0, 0.0), S0)
        elapsed += time.perf_counter() - t0
        its, e = 0, None
        while its < max_iter:
            t0 = time.perf_counter()
            m.run_iterations(force, lat, chunk)
            elapsed += time.perf_counter() - t0
            its += chunk
            e = force.compute_energy(lat)[0] / lat.n_active()
            if e <= target:
                break
        out[mode] = dict(seconds=elapsed, iterations=its, energy_per_site=e, reached=bool(e <= target))
        print("energy of the random start (%.10f) reached from the %s start: %d iterations, %.2f s"
              % (target, mode, its, elapsed), file=sys.stderr, flush=True)
        del lat
    return out


def warm_start_heldout(qmin, L=128, seeds=(12, 13), random_seed=4321, fc=1e-8, weights=WEIGHTS):
    """The learned warm start on held-out periodic L^3 geometries with 2-6
    spheres (qmin.warmstart.heldout_geometry; 5 and 6 spheres for seeds 12 and
    13): FIRE to forceMax < fc from a seeded random start, from a uniform
    director along a body diagonal, and from the network's prediction along the
    same diagonal (time includes the prediction), all with FP32 state and FP64
    reductions."""
    from qmin import warmstart as ws
    net = ws.WarmStartNet(seed=0)
    net.load(weights)
    A, B, C, S0 = phase_constants()
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    d = ws.diagonal_director()
    out = {}
    for seed in seeds:
        spheres = ws.heldout_geometry(L, seed)
        res = {"spheres": len(spheres)}
        for mode in ("random", "uniform_diagonal", "network_diagonal"):
            lat = qmin.Lattice(L, L, L, precision=qmin.QM_FP32)
            if mode == "random":
                lat.random_init(random_seed, S0)
            else:
                lat.uniform_init(d, S0)
            for c, r, t, W in spheres:
                lat.sphere(c, r, t, W, S0)
            t0 = time.perf_counter()
            if mode.startswith("network"):
                net.predict(lat, d, S0)
            st = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=fc, max_iter=400000).minimize(force, lat)
            t1 = time.perf_counter()
            e, _ = force.compute_energy(lat)
            res[mode] = dict(seconds=t1 - t0, iterations=st.iterations, forceMax=st.forceMax,
                             converged=bool(st.converged), energy_per_site=e / lat.n_active())
            print("held-out %d^3 seed %d, %s start: %d iterations, %.2f s" % (L, seed, mode, st.iterations, t1 - t0),
                  file=sys.stderr, flush=True)
            del lat
        er = res["random"]["energy_per_site"]
        for mode in ("uniform_diagonal", "network_diagonal"):
            res[mode]["speedup_vs_random"] = res["random"]["seconds"] / res[mode]["seconds"]
            res[mode]["energy_rel_diff_vs_random"] = (res[mode]["energy_per_site"] - er) / abs(er)
        res["network_diagonal"]["speedup_vs_uniform"] = res["uniform_diagonal"]["seconds"] / \
            res["network_diagonal"]["seconds"]
        out["seed_%d" % seed] = res
    return out


def _pair(fc, t64, tmx):
    return {"forceMax_below": fc, "fp64": t64, "mixed_fp32_then_fp64_polish": tmx,
            "mixed_speedup": t64["seconds"] / tmx["seconds"],
            "energy_rel_diff": abs(tmx["energy_per_site"] - t64["energy_per_site"]) / abs(t64["energy_per_site"])}


def compare_modes(qmin, L, fc=1e-8, tight=(1e-10, 1e-11), switch_fc=MIXED_SWITCH):
    """FP64 against mixed precision to forceMax < fc (FP32 until fc, then the
    FP64 polish) and, as one continued run per mode, to each tighter tolerance
    (FP32 until switch_fc near its rounding floor).  Returns the fc comparison,
    with the defect ring of the FP64 state, and the tight ones."""
    fp64 = tolerance_chain(qmin, L, "fp64", (fc,) + tuple(tight), ring=True)
    mixed = time_to_tolerance(qmin, L, "mixed", fc)
    first = _pair(fc, fp64[0], mixed)
    first["defect_ring"] = fp64[0].pop("defect_ring")
    first["fp64_max_site_force"] = fp64[0].pop("max_site_force")
    tight_out = {}
    if tight:
        mixed_tight = tolerance_chain(qmin, L, "mixed", tuple(tight), switch_fc)
        for k, t in enumerate(tight):
            tight_out["%g" % t] = _pair(t, fp64[k + 1], mixed_tight[k])
        # how far the FP64 energy itself still moves between the tight tolerances
        if len(tight) > 1:
            e = [fp64[k + 1]["energy_per_site"] for k in range(len(tight))]
            tight_out["fp64_energy_rel_change_%g_to_%g" % (tight[0], tight[-1])] = abs(e[-1] - e[0]) / abs(e[0])
    return first, tight_out


def ensemble_throughput(qmin, count=1024, fc=1e-7, max_iter=10000, sequential_sample=32, graph_iterations=32):
    """The ensemble workload (seeded uneven lattices with spheres, L1 only, FP32
    state with FP64 reductions): the tile autotuner picks the batched pass's
    shape (outside the timed region), then FIRE on all lattices at once, batched
    launches replayed from a hipGraph; the first lattices are also minimized one
    by one for comparison.  dt_max is 40 dt: at openQmin's default 100 dt FIRE diverges on
    stiff planar-anchored spheres (the reference algorithm does too)."""
    from qmin import workloads
    A, B, C, S0 = workloads.phase_constants()
    specs = workloads.ensemble_specs(count)
    t0 = time.perf_counter()
    lats = [workloads.build_lattice(qmin, sp, qmin.QM_FP32, S0) for sp in specs]
    build_s = time.perf_counter() - t0
    sites = [l.n for l in lats]
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=qmin.SEMANTICS_REF_GPU)
    params = dict(dt=5e-4, dt_max=2e-2, fc=fc, max_iter=max_iter)
    ens = qmin.Ensemble(lats)
    ta = time.perf_counter()
    tile, zchunk, pass_ms = ens.autotune(force)
    t1 = time.perf_counter()
    its, fms = ens.fire_minimize(qmin.FIRE(**params), force, graph_iterations)
    t2 = time.perf_counter()
    updates = sum(n * int(i) for n, i in zip(sites, its))
    seq = [workloads.build_lattice(qmin, sp, qmin.QM_FP32, S0) for sp in specs[:sequential_sample]]
    t3 = time.perf_counter()
    seq_its = [qmin.FIRE(**params).minimize(force, l).iterations for l in seq]
    t4 = time.perf_counter()
    seq_updates = sum(n * i for n, i in zip(sites, seq_its))
    print("ensemble: %d lattices, %.3e site-updates/s" % (count, updates / (t2 - t1)), file=sys.stderr, flush=True)
    return {"lattices": count, "sites": int(sum(sites)), "dtype": "f32 (FP64 reductions)",
            "forceMax_below": fc, "max_iterations": max_iter, "seconds": t2 - t1, "build_seconds": build_s,
            "site_updates_per_s": updates / (t2 - t1), "converged": int(sum(1 for f in fms if f <= fc)),
            "autotune": {"tile": "per lattice" if tile == qmin.TILE_PER_LATTICE else ["32x8", "64x4", "32x16", "64x8"][tile],
                         "zchunk": zchunk, "pass_ms": pass_ms, "seconds": t1 - ta},
            "iterations_min_mean_max": [int(min(its)), float(sum(its)) / count, int(max(its))],
            "one_by_one": {"lattices": sequential_sample, "site_updates_per_s": seq_updates / (t4 - t3),
                           "same_iterations": [int(i) for i in its[:sequential_sample]] == seq_its}}


def full_elasticity_iteration_times(qmin, L=384, iterations=10):
    """Per-iteration wall time of FIRE, Adam, Nesterov and online L-BFGS (memory
    8) on the full-elasticity workload (qmin.workloads.full_elasticity: L1..L4,
    L6, E field, four spheres) in FP64, each after two warm-up iterations, on one
    lattice in turn (a bounded leg of the default run; --full-elasticity runs each
    to tolerance)."""
    from qmin import workloads
    lat, f = workloads.full_elasticity(qmin, L, qmin.QM_FP64)
    runs = (("FIRE", qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=1e-12, max_iter=10 ** 9)),
            ("Adam", qmin.Adam(dt=1e-4, fc=1e-12, max_iter=10 ** 9)),
            ("Nesterov", qmin.Nesterov(dt=1e-4, mu=0.01, fc=1e-12, max_iter=10 ** 9)),
            ("LoLBFGS(m=8)", qmin.LoLBFGS(m=8, dt=1e-4, fc=1e-12, max_iter=10 ** 9)))
    out = {"size": L, "dtype": "f64", "iterations": iterations}
    for name, m in runs:
        if name == "FIRE":
            m.run_iterations(f, lat, 2)
            t0 = time.perf_counter()
            m.run_iterations(f, lat, iterations)
        else:
            m.set_max_iterations(2)
            m.minimize(f, lat)
            m.set_max_iterations(2 + iterations)
            t0 = time.perf_counter()
            m.minimize(f, lat)
        ms = (time.perf_counter() - t0) / iterations * 1e3
        out[name] = {"ms_per_iteration": ms, "site_updates_per_s": lat.n / ms * 1e3}
        print("full elasticity %d^3 FP64, %s: %.3f ms per iteration" % (L, name, ms), file=sys.stderr, flush=True)
    del lat
    return out


# Minimizers of the full-elasticity workload: (name, class, parameters, iteration
# cap of each phase).  Step sizes: Nesterov's 1e-2 is the largest of {1e-4, 1e-3,
# 1e-2, 2e-2, 5e-2} that stays finite on this workload at 128^3; FIRE's dt_max is
# the ensemble's.  Online L-BFGS (no line search, energyMinimizerLoLBFGS.cpp:104-187)
# converges at 128^3 with dt = 5e-4 but diverges at 384^3 for dt >= 1e-4 (after
# about 780 iterations at 1e-4, 100 at 2e-4); at 5e-5 it stays finite and runs a
# fixed number of iterations.  With the reference's scheduledMomentum
# (energyMinimizerLoLBFGS.cpp:205-208: step tau / (tau + iteration) x dt, tau = 10)
# it runs at dt = 1e-4 with a step that shrinks as 1/iteration.  Adam keeps steps
# of size ~dt however small the force (m/sqrt(s) does not shrink with F): at a
# constant dt it stalls near forceMax ~ 1e-6 and also runs a fixed number of
# iterations.
FULL_ELASTICITY_MINIMIZERS = (
    ("FIRE", "FIRE", dict(dt=5e-4, dt_max=2e-2), 20000),
    ("Nesterov", "Nesterov", dict(dt=1e-2, mu=0.01), 20000),
    ("LoLBFGS(m=8)", "LoLBFGS", dict(m=8, dt=5e-5), 500),
    ("LoLBFGS(m=8, scheduledMomentum tau=10)", "LoLBFGS", dict(m=8, dt=1e-4, tau=10.0, scheduled=True), 1000),
    ("Adam", "Adam", dict(dt=1e-3), 1000),
)


def full_elasticity(qmin, L=384, fc=1e-8):
    """Each minimizer from the same seeded start of the full-elasticity workload
    (qmin.workloads.full_elasticity: L1..L4, L6, E field, four spheres) until
    forceMax < fc or its iteration cap, in FP64 and mixed: FP32 state until
    forceMax < fc, then the state is converted on the device and FP64 polishes
    to fc (FIRE continues with its dt and alpha; the other minimizers restart
    their history in FP64, as a new minimizer on the converted lattice)."""
    from qmin import workloads
    out = {"size": L, "forceMax_below": fc}
    for name, cls, params, cap in FULL_ELASTICITY_MINIMIZERS:
        res = {"parameters": dict(params), "iteration_cap": cap}
        for mode in ("fp64", "mixed"):
            res[mode] = _full_elasticity_run(qmin, workloads, L, fc, cls, params, cap, mode)
            r = res[mode]
            print("full elasticity %d^3 %s %s: %s" % (L, name, mode, r.get("error") or "%d iterations, %.2f s, forceMax "
                                                      "%.3g" % (r["iterations"], r["seconds"], r["forceMax"])),
                  file=sys.stderr, flush=True)
        if "error" not in res["fp64"] and "error" not in res["mixed"]:
            res["mixed_speedup"] = res["fp64"]["seconds"] / res["mixed"]["seconds"]
            res["energy_rel_diff_mixed_vs_fp64"] = abs(res["mixed"]["energy_per_site"] -
                                                       res["fp64"]["energy_per_site"]) / abs(
                res["fp64"]["energy_per_site"])
        out[name] = res
    return out


def _full_elasticity_run(qmin, workloads, L, fc, cls, params, cap, mode):
    """One minimizer in one mode on a fresh full-elasticity lattice."""
    lat, force = workloads.full_elasticity(qmin, L, qmin.QM_FP64 if mode == "fp64" else qmin.QM_FP32)
    m = getattr(qmin, cls)(fc=fc, max_iter=cap, **params)
    t0 = time.perf_counter()
    try:
        st = m.minimize(force, lat)
    except qmin.QminError as e:
        st = m.state()
        return {"error": str(e), "iterations": st.iterations if cls == "FIRE" else st[0],
                "seconds": time.perf_counter() - t0}
    t1 = time.perf_counter()
    its, fmax = (st.iterations, st.forceMax) if cls == "FIRE" else st
    r = {"fp32_iterations": its, "fp32_seconds": t1 - t0} if mode == "mixed" else {}
    if mode == "mixed" and fmax < fc:
        lat.set_precision(qmin.QM_FP64)
        if cls == "FIRE":
            m.set_max_iterations(its + cap)
            st = m.minimize(force, lat)
            its, fmax = st.iterations, st.forceMax
        else:
            m2 = getattr(qmin, cls)(fc=fc, max_iter=cap, **params)
            p_its, fmax = m2.minimize(force, lat)
            its += p_its
    t2 = time.perf_counter()
    e, _ = force.compute_energy(lat)
    r.update(iterations=its, seconds=t2 - t0, forceMax=fmax, converged=bool(fmax < fc),
             energy_per_site=e / lat.n_active(), ms_per_iteration=1e3 * (t2 - t0) / max(its, 1))
    if mode == "mixed":
        # no polish when the FP32 phase stopped at its cap
        r.update(polish_iterations=its - r["fp32_iterations"], polish_seconds=t2 - t1)
    return r


def measured_traffic(kernel, size, precision):
    """HBM bytes per dispatch of the FIRE lattice kernel from the committed
    rocprofv3 counter summary of this workload, when it covers this kernel,
    size and precision."""
    path = os.path.join(REPO, "profiles", "%s_fire_%s_%d.traffic.json" % (kernel, precision, size))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d["read_bytes"] + d["write_bytes"]


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(L=256, threads_all=None, reps=3):
    """The CPU oracle (a restatement of the reference CPU path, which is
    single-threaded) on this workload at full size: one FIRE iteration is the
    time of fire(max_iter=1) (one force evaluation and one iteration) minus one
    force evaluation, each the median of `reps` runs (about 45 s on one
    core at 256^3); on one core, and with
    OpenMP on every host core this process may use (the reference has no
    threading; that row is for comparison)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po
    A, B, C, S0 = phase_constants()
    o = po.OracleSystem(L, L, L)
    o.random_init(1234, S0)
    o.wall(2, 0, 1, 5.3, S0)
    o.wall(2, L - 1, 1, 5.3, S0)
    o.sphere((L / 2.0, L / 2.0, L / 2.0), 20.0 * L / 256.0, 0, 5.3, S0)
    o.set_params(A, B, C, 2.32, 0.5, 0.3, semantics=1)
    q0 = o.q()
    all_cores = threads_all or host_cores()
    rows = {}
    t_start = time.perf_counter()
    for threads in (1, all_cores):
        o.threads(threads)
        t_force, t_fire = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            o.forces()
            t_force.append(time.perf_counter() - t0)
            o.set_q(q0)
            t0 = time.perf_counter()
            o.fire(max_iter=1)
            t_fire.append(time.perf_counter() - t0)
        per_it = sorted(t_fire)[reps // 2] - sorted(t_force)[reps // 2]
        rows[threads] = dict(value=L ** 3 / per_it, s_per_iteration=per_it)
    seconds = time.perf_counter() - t_start
    one, many = rows[1], rows[all_cores]
    return dict(value=one["value"], unit="site-updates/s", cores=1, kind="port",
                s_per_iteration=one["s_per_iteration"], host_cores=host_cores(),
                all_cores={"value": many["value"], "cores": all_cores, "s_per_iteration": many["s_per_iteration"]},
                sample="the bench workload at %d^3 (walls, sphere, L1-L3, FP64): oracle/qmin_oracle.cpp, one FIRE "
                       "iteration = fire(1 iteration) - one force evaluation, median of %d runs each, on 1 core "
                       "(the reference CPU path is single-threaded) and with OpenMP on %d cores; %.1f s in all"
                       % (L, reps, all_cores, seconds))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--zchunk", type=int, default=0, help="fixed z chunk of the fused pass (0 = automatic)")
    ap.add_argument("--no-ttt", action="store_true", help="skip the time-to-tolerance runs")
    ap.add_argument("--ttt-fc", type=float, default=1e-8, help="forceMax tolerance of the time-to-tolerance runs")
    ap.add_argument("--ensemble", action="store_true",
                    help="also measure the 1024-lattice ensemble workload (about 2 minutes)")
    ap.add_argument("--full-elasticity", action="store_true",
                    help="also run every minimizer to tolerance on the 384^3 full-elasticity workload in FP64 and "
                         "mixed precision (several minutes)")
    ap.add_argument("--no-energy-check", action="store_true",
                    help="skip the FP64-vs-mixed comparisons at forceMax < 1e-10 and 1e-11")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's hot path runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if not args.no_ttt:
        # the warm start uses torch for buffers: load its HIP runtime before the
        # library so that one runtime serves both
        import torch  # noqa: F401
    import qmin
    qmin.library()
    precision = qmin.QM_FP64 if args.precision == "fp64" else qmin.QM_FP32
    r = gpu_throughput(qmin, args.size, precision, args.steps, args.warmup, args.zchunk)
    value = r["n"] * args.steps / r["wall"]
    bps = BYTES_PER_SITE[precision]
    achieved = bps * r["n"] / (r["march_ms"] * 1e-3) / 1e9
    if args.size < 128:
        kernel, kernel_desc, state_bytes = "march_kernel", "fused FIRE lattice pass (march_kernel, 32x8 / 64x4 " \
                                                           "tile)", bps
    elif precision == qmin.QM_FP64:
        # the FIRE pass that keeps only Q and v: it reads and writes Q and v (and reads the site class), and
        # recomputes the previous force on a one-site halo instead of storing F
        kernel, state_bytes = "fire_stream_kernel", 161.0
        kernel_desc = "FIRE pass keeping Q and v only (fire_stream_kernel, 32x16 tile)"
    else:
        kernel, kernel_desc, state_bytes = "march_wide_kernel", "fused FIRE lattice pass (march_wide_kernel, 64x8 " \
                                                                "tile)", bps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == qmin.QM_FP64 else "f32 (FP64 reductions)",
        "data": "synthetic: seeded random directors, walls and sphere generated in-process",
        "config": {"workload": "FIRE, %d^3, planar walls z=0/z=%d, homeotropic sphere r=%g, L1=2.32 L2=0.5 L3=0.3"
                               % (args.size, args.size - 1, 20.0 * args.size / 256.0),
                   "sites": r["n"], "boundary_sites": r["n_boundary"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(kernel, args.size, args.precision),
                     "kernel": kernel_desc, "kernel_ms": r["march_ms"], "boundary_kernel_ms": r["boundary_ms"],
                     "bytes_per_site": bps, "kernel_state_bytes_per_site": state_bytes},
        "iteration_hbm_frac": value * bps / (PEAK_HBM_GBS * 1e9),
        "forceMax_after": r["forceMax"],
    }
    if not args.no_ttt:
        tight = () if args.no_energy_check else (1e-10, 1e-11)
        out["time_to_tolerance"], checks = compare_modes(qmin, args.size, args.ttt_fc, tight)
        if checks:
            # forceMax < 1e-8 (= |F|_2 / N) leaves the slow elastic modes unconverged by ~1e-6 in E/N whatever
            # the precision; the soft azimuthal mode of the director between the planar walls still moves the
            # energy by a few 1e-9 between 1e-10 and 1e-11
            out["converged_energy_check"] = checks
        ws = warm_start(qmin, args.size, args.ttt_fc)
        t_random = out["time_to_tolerance"]["fp64"]
        for v in ws.values():
            v["speedup_vs_random_fp64"] = t_random["seconds"] / v["seconds"]
            v["energy_rel_diff_vs_random_fp64"] = (v["energy_per_site"] - t_random["energy_per_site"]) / abs(
                t_random["energy_per_site"])
        # the same comparison at matched energy: the warm starts end in lower minima than the random start,
        # so the time until they reach the random start's final energy is the like-for-like figure
        em = warm_start_energy_matched(qmin, args.size, t_random["energy_per_site"])
        for mode in ("network", "uniform_director"):
            em[mode]["speedup_vs_random_fp64"] = t_random["seconds"] / em[mode]["seconds"]
        ws["energy_matched"] = em
        out["time_to_tolerance"]["warm_start"] = ws
        out["time_to_tolerance"]["warm_start_heldout_128"] = warm_start_heldout(qmin)
    if args.ensemble:
        out["ensemble"] = ensemble_throughput(qmin)
    elif not args.no_ttt:
        # bounded legs of the two other workloads, so that the default run records them too
        out["ensemble_256"] = ensemble_throughput(qmin, count=256, max_iter=512, sequential_sample=8)
    if args.full_elasticity:
        out["full_elasticity"] = full_elasticity(qmin, fc=args.ttt_fc)
    elif not args.no_ttt:
        out["full_elasticity_iteration_times"] = full_elasticity_iteration_times(qmin)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
