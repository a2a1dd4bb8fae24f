"""Throughput of the spectral_rs hot path on one MI355X.

Workload (N=1): the plain FP32 unmagnetized shock on a 4096 x 1024 grid,
32 particles per cell per species (ions m=16, electrons m=1; delta = 16 gives
2 x 133,169,152 = 266,338,304 particles), gamma0 = 10, c = 10 cells/omega_pe
with dt = 0.045 (0.45 cells/step, 10-cell skin depth), 4 binomial passes.
Particles are generated on the GPU (positions exactly as the reference,
momenta from a seeded counter-based RNG); they are resident in HBM before the
timed region.  One step = one iteration of the reference's run() loop:
zero, push+move+deposit of both species, ghost fold, filter, 7 forward and
6 inverse 2D FFTs with the spectral Maxwell update.

Prints one JSON line (rank 0) with the metric, a roofline object for the
dominant kernel (the fused particle pass) and a CPU baseline: the oracle's
C++ restatement of the reference CPU path, timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

from spectral_rs_amd.workloads import C2, SPECIES, ensemble_runs, particles  # noqa: E402


def crop_error():
    """The FP32 GPU path on the 512 x 128 slab of this shock (32 ppc per
    species, 3,932,160 particles, 1000 steps) against the FP64 (dprec)
    oracle's snapshots at steps 100, 500 and 1000 (tests/golden, written by
    tests/golden/make_shock_snapshots.py), next to the FP32 oracle's own
    errors.  Untimed; None when the snapshots are absent."""
    import numpy as np

    from spectral_rs_amd import capi
    from spectral_rs_amd.analysis import HORIZON_CASES, SHOCK_PHYS, horizon_errors, inputs_digest, shock_inputs

    path = os.path.join(ROOT, "tests", "golden", "shock_oracle_snapshots.npz")
    if not os.path.exists(path):
        return None
    k = HORIZON_CASES["c2_crop"]
    sp = shock_inputs("c2_crop")
    with np.load(path, allow_pickle=False) as z:
        if inputs_digest(sp) != str(z["c2_crop_inputs_sha256"]):
            return None
    g = capi.GpuSim(prec=4, sort_interval=20, left_edge=k["left_edge"], size_x=k["size_x"], size_y=k["size_y"],
                    delta=k["delta"], dens=k["dens"], **SHOCK_PHYS)
    for q, m, a in sp:
        g.add_species(q, m, **a)
    g.prestep()
    res = horizon_errors(g, "c2_crop", path)
    g.close()
    return {"workload": "512x128 slab of this shock, 32 ppc/species, identical seeded inputs on both sides",
            "reference": "FP64 (dprec) oracle",
            "metrics": "relative L2: y-averaged density profile, interior Ex Ey Bz, x power spectra of Ex Bz",
            "steps": {str(r["step"]): {"gpu_fp32": r["errors"], "oracle_fp32": r["oracle_fp32_errors"]}
                      for r in res}}


def cpu_baseline(seconds_hint=15.0):
    """The oracle (restated reference CPU path, f32, OpenMP like rayon) on a
    512 x 128 crop of the same physics with 32 ppc per species, on the host's
    threads and again on one thread."""
    import numpy as np

    from oracle import oracle as O
    from spectral_rs_amd.inputs import shock_species

    cfg = dict(C2, size_x=512, size_y=128)
    r = O.RefSim(prec=4, **cfg)
    for k, (q, m, vth) in enumerate(SPECIES):
        r.add_species(q, m, **shock_species(cfg["size_x"], cfg["size_y"], cfg["delta"], cfg["dens"],
                                            cfg["c"], cfg["gamma_inj"], vth, seed=k, ft=np.float32))
    r.prestep()
    t0 = time.perf_counter()
    r.step(1)
    one = time.perf_counter() - t0
    steps = max(2, int(seconds_hint / max(one, 1e-3)))
    t0 = time.perf_counter()
    r.step(steps)
    el = time.perf_counter() - t0
    n = particles(cfg)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    thr = int(os.environ.get("OMP_NUM_THREADS", cores))
    # one thread: the reference's rayon regions serialised
    O.set_threads(1)
    steps1 = max(2, steps // 4)
    t0 = time.perf_counter()
    r.step(steps1)
    el1 = time.perf_counter() - t0
    O.set_threads(thr)
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": thr, "kind": "port",
            "sample": f"oracle f32, 512x128 crop, 32 ppc/species ({n} particles), {steps} steps, "
                      f"{el:.1f} s; serial mover/deposit, threaded push/filter/FFT as the reference",
            "single_thread": {"value": n * steps1 / el1, "cores": 1, "sample": f"the same, {steps1} more steps, "
                                                                              f"{el1:.1f} s"}}


def bench_ensemble(args):
    """64 independent runs of uneven size stepped together: one hipGraph per
    `unit` steps of every member in `branches` concurrent branches, tunings
    from the autotuner.  Also times the same ensemble eagerly and with default
    tunings."""
    from spectral_rs_amd import capi

    steps = args.steps if args.steps_given else 1000
    runs = ensemble_runs()
    sims = []
    n = 0
    cells = 0
    for r in runs:
        kw = {k: v for k, v in r.items() if k != "temp"}
        s = capi.GpuSim(prec=4, sort_interval=args.sort_interval, left_edge=capi.SP_LEFT_REFLECT, **kw)
        for q, m in [(1.0, 16.0), (-1.0, 1.0)]:
            s.init_species(q, m, (r["temp"] / m) ** 0.5, seed=0)
        n += 2 * r["dens"] * (r["size_x"] - 2 * r["delta"]) * r["size_y"]
        cells += (r["size_x"] + 2) * (r["size_y"] + 2)
        sims.append(s)
    print(f"bench: ensemble of {len(sims)} runs, {n} particles, {cells} cells", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tunings = [s.autotune(trials=3, use_cache=True) for s in sims]
    t_tune = time.perf_counter() - t0
    shapes = {}
    for r, t in zip(runs, tunings):
        shapes[f"{r['size_x']}x{r['size_y']}/{r['dens']}ppc"] = t
    print(f"bench: autotuned {len(shapes)} distinct configurations in {t_tune:.1f} s", file=sys.stderr, flush=True)
    for s in sims:
        s.prestep()
    ens = capi.Ensemble(sims, branches=args.branches, unit=args.unit)
    ens.step(max(args.warmup, 2))
    ens.sync()

    def timed(k):
        ens.sync()
        t0 = time.perf_counter()
        ens.step(k)
        ens.sync()
        return time.perf_counter() - t0

    cmp_steps = 100
    el = timed(steps)
    ens.set_graph(False)
    el_eager = timed(cmp_steps)
    ens.set_graph(True)
    ens.close()
    for s in sims:
        s.set_tuning(chunk=2048, tile_w_log=3, tile_h_log=3, fft_rows=1, fft_cols=8, fft_radix_x=8, fft_radix_y=8)
    ens = capi.Ensemble(sims, branches=args.branches, unit=args.unit)
    ens.step(2)
    el_default = timed(cmp_steps)
    st = ens.stats()
    ens.close()
    for s in sims:
        s.close()
    value = n * steps / el
    ms = el * 1e3 / steps
    # particle-pass algorithmic bytes of the whole ensemble per step over the
    # whole step's time: a lower bound on the rate of the passes themselves
    alg = 48.0 * n + 2 * 44.0 * cells
    out = {
        "metric": "particle-steps/s on 1 MI355X (batched ensemble)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 2),
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-generated shock initial states, seed 0)",
        "config": {"workload": "ensemble of 64 independent FP32 shock runs, grids {128..1024}x{32..256}, "
                               "ppc {4..32}, gamma0 in [2,20], T log-uniform in [1e-5,1e-3], 4 binomial passes; "
                               f"one hipGraph per {args.unit} steps of all runs, {args.branches} branches, "
                               "autotuned launch shapes",
                   "runs": len(runs), "particles": n, "cells": cells},
        "modes_ms_per_step": {"graph_tuned": ms, "eager_tuned": el_eager * 1e3 / cmp_steps,
                              "graph_default_tuning": el_default * 1e3 / cmp_steps},
        "autotune_s": t_tune,
        "tunings": shapes,
        "graphs": st,
        "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "particle passes of all runs, algorithmic bytes per whole step"},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-crop", action="store_true", help="skip the 512x128 comparison with the FP64 oracle")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--size", default=None, help="override grid, e.g. 1024x256 (not the metric workload)")
    ap.add_argument("--sort-interval", type=int, default=20, help="steps between tile sorts of the particles")
    ap.add_argument("--sort-policy", default="3,40,0.03",
                    help="adaptive tile-sort schedule min,max,rise (sp_set_sort_policy), or 'fixed' for every "
                         "--sort-interval steps")
    ap.add_argument("--trace", type=int, default=0, help="print phase times every N steps (diagnostic)")
    ap.add_argument("--autotune", action="store_true", help="shock workload: autotune the launch shapes first")
    ap.add_argument("--branches", type=int, default=4, help="ensemble workload: concurrent branches")
    ap.add_argument("--unit", type=int, default=2, help="ensemble workload: steps per captured graph")
    ap.add_argument("--workload", choices=["shock", "ensemble"], default="shock",
                    help="shock: the 4096x1024 run (the metric); ensemble: 64 uneven runs stepped together")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only; multi-GPU runs are not supported", file=sys.stderr)
        sys.exit(2)
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 2000
    if args.workload == "ensemble":
        bench_ensemble(args)
        return

    from spectral_rs_amd import capi

    cfg = dict(C2)
    if args.size:
        sx, sy = (int(v) for v in args.size.lower().split("x"))
        cfg.update(size_x=sx, size_y=sy)
    n = particles(cfg)
    cells = (cfg["size_x"] + 2) * (cfg["size_y"] + 2)
    print(f"bench: {cfg['size_x']}x{cfg['size_y']}, {n} particles, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    # The reference has no left boundary and asserts when a particle reaches
    # ix <= delta (src/prtls.rs:163-170), which the rear edge of the slab does
    # in long runs; the bench uses the mirrored-wall left edge (SP_LEFT_REFLECT).
    sim = capi.GpuSim(prec=4, sort_interval=args.sort_interval, left_edge=capi.SP_LEFT_REFLECT, **cfg)
    policy = None
    if args.sort_policy != "fixed":
        lo, hi, rise = args.sort_policy.split(",")
        policy = (int(lo), int(hi), float(rise))
        sim.set_sort_policy(*policy)
    for k, (q, m, vth) in enumerate(SPECIES):
        sim.init_species(q, m, vth, seed=0)
    tuning = None
    if args.autotune:
        t0 = time.perf_counter()
        tuning = sim.autotune(trials=3, use_cache=False)
        print(f"bench: autotuned in {time.perf_counter() - t0:.1f} s: {tuning}", file=sys.stderr, flush=True)
    sim.prestep()
    sim.step(args.warmup)
    sim.sync()
    sim.set_timing(True)
    sim.timers(reset=True)
    sorts0 = sim.diag()["sorts"]
    t0 = time.perf_counter()
    if args.trace:  # phase times every `trace` steps (adds a sync per block of steps)
        tm = {"dep": 0.0, "solve": 0.0, "push": 0.0, "particles": 0.0}
        done = 0
        while done < args.steps:
            k = min(args.trace, args.steps - done)
            sim.step(k)
            sim.sync()
            t = sim.timers(reset=True)
            done += k
            print(f"trace: steps {done - k}-{done}: " + " ".join(f"{n} {v / k:.3f} ms" for n, v in t.items()),
                  file=sys.stderr, flush=True)
            for n_ in tm:
                tm[n_] += t[n_]
    else:
        sim.step(args.steps)
        sim.sync()
    el = time.perf_counter() - t0
    if not args.trace:
        tm = sim.timers(reset=True)
    nsorts = sim.diag()["sorts"] - sorts0
    value = n * args.steps / el
    print(f"bench: {el:.3f} s, {value:.4g} particle-steps/s, phases ms {tm}", file=sys.stderr, flush=True)

    # Fused particle pass (push + move + deposit), one launch per species per
    # step.  Algorithmic bytes of one launch = 48 B per particle of the species
    # (24 B state read + 24 B written) + 44 B per padded cell (6 field grids
    # gathered, 5 deposit grids written); its duration is the phase time of
    # the pass (HIP events on the library's stream) over the launches.
    launches = len(SPECIES)
    part_ms = tm["particles"] / args.steps / launches
    traffic_alg = 48.0 * n / launches + 44.0 * cells
    achieved = traffic_alg / (part_ms * 1e-3) / 1e9
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE on this
    # workload (calibrated against kernels of known traffic), when that
    # measurement exists.
    traffic = None
    tf = os.path.join(ROOT, "profiles", "c2_particles_traffic.json")
    if not args.size and os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        traffic = 1024.0 * (t["fetch_kib_per_species_launch"] * t["fetch_factor"]
                            + t["write_kib_per_species_launch"] * t["write_factor"])
    out = {
        "metric": "particle-steps/s on 1 MI355X",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU-generated shock initial state, seed 0)",
        "config": {"workload": f"plain FP32 unmagnetized shock {cfg['size_x']}x{cfg['size_y']}, 32 ppc/species, "
                               f"delta 16, gamma0 10, mass ratio 16, 4 binomial passes",
                   "particles": n, "grid": [cfg["size_x"], cfg["size_y"]],
                   "tuning": tuning or sim.tuning(),
                   "sort_schedule": ({"adaptive": {"min": policy[0], "max": policy[1], "rise": policy[2]}} if policy
                                     else {"every": args.sort_interval})},
        "phases_ms_per_step": {k: v / args.steps for k, v in tm.items()},
        "window": (f"{args.steps} steps: {nsorts} tile sorts ("
                   + (f"adaptive schedule: {policy[0]} to {policy[1]} steps apart, sooner when the share of "
                      f"particles on the particle pass's general path rose by {policy[2]}" if policy else
                      f"one per {args.sort_interval} steps") +
                   "), and one stand-alone Boris push at the end (the last step's push, otherwise fused into "
                   "the next step's particle pass); the roofline duration is the fused particle pass with its "
                   "share of the species-density fold that follows both passes (k_fold_species_density, ~23 us)"),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_particles<float, true> (fused push + move + deposit, one species per launch)",
                     "launches_per_step": launches, "bytes_per_launch_algorithmic": traffic_alg,
                     "ms_per_launch": part_ms},
    }
    sim.close()
    if not args.size and not args.no_crop:
        ce = crop_error()
        if ce is not None:
            out["fp64_oracle_error"] = ce
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
