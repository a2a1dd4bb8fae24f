"""Writes tests/golden/shock_oracle_snapshots.npz: the oracle's shock runs that
the GPU is judged against over long horizons (test infrastructure; runs the
CPU oracle, ~15 minutes on 8 cores).

    python tests/golden/make_shock_snapshots.py

Cases (reference units: c = 10, dt = 0.045, so 0.45 cells per step and a
10-cell skin depth; ions q = +1, m = 16, vth 0.0025, then electrons q = -1,
m = 1, vth 0.01; gamma0 = 10; 4 binomial passes; inputs from
tests.helpers.shock_species with seeds 0 and 1, uploaded identically to every
run):

* c2_crop: the 512 x 128 slab of the 4096 x 1024 shock, delta 16, 32 ppc per
  species (3,932,160 particles), snapshots at steps 100, 500 and 1000.  The
  slab's rear edge reaches delta at ~1070 steps; the run uses the mirrored
  left edge (left_edge = 1), which changes nothing before that.
* c1: 256 x 64, delta 8, 16 ppc per species (491,520 particles), snapshots
  at steps 100 and 300 and the last step before the reference's left-edge
  assert (src/prtls.rs:163-170) fires, the step of which is recorded.

For each snapshot: the FP64 (`dprec`) oracle's y-averaged density profile,
its interior Ex, Ey, Bz (stored as float32), the x power spectra of Ex and Bz
and the energies; and the FP32 oracle's own errors against it -- the yardstick
for the GPU's FP32 errors, which come from the deposit's summation order and
the FFT instead of the serial loop.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from spectral_rs_amd.analysis import (HORIZON_CASES, SHOCK_PHYS, inputs_digest, shock_inputs,  # noqa: E402
                                     shock_metrics, snapshot_errors)
from tests.helpers import energies  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shock_oracle_snapshots.npz")
CASES = HORIZON_CASES
PHYS = SHOCK_PHYS


def inputs(case):
    return shock_inputs(case)


def digest(sp):
    return inputs_digest(sp)


def oracle_run(case, prec, sp):
    k = CASES[case]
    kw = dict(PHYS, size_x=k["size_x"], size_y=k["size_y"], delta=k["delta"], dens=k["dens"],
              left_edge=k["left_edge"])
    s = O.RefSim(prec=prec, **kw)
    for q, m, a in sp:
        # gamma (psa) from the momenta in the run's own precision, as
        # initialize_velocities does (src/prtls.rs:248-249): the float32 psa
        # of the inputs would put a float32 rounding into the dprec run
        a = {kk: v for kk, v in a.items() if kk != "psa"}
        s.add_species(q, m, **({kk: (v.astype(np.float64) if v.dtype == np.float32 else v) for kk, v in a.items()}
                               if prec == 8 else a))
    s.prestep()
    return s


def snapshot(s, k):
    g = {n: s.grid(n)[1:-1, 1:-1].astype(np.float32) for n in ("ex", "ey", "bz")}
    m = shock_metrics(s, k["dens"])
    we, wb, kin, _ = energies(s, [16.0, 1.0], k["dens"], PHYS["c"])
    return g, m, np.array([we, wb] + list(kin))


def main():
    O.set_threads(os.cpu_count())
    out = {}
    for case, k in CASES.items():
        sp = inputs(case)
        out[f"{case}_inputs_sha256"] = np.array(digest(sp))
        runs = {8: oracle_run(case, 8, sp), 4: oracle_run(case, 4, sp)}
        steps = list(k["steps"])
        t = 0
        tic = time.time()
        snaps = {8: [], 4: []}
        for target in steps:
            for prec, s in runs.items():
                s.step(target - t)
                snaps[prec].append(snapshot(s, k))
            t = target
            print(f"{case}: step {t} ({time.time() - tic:.0f} s)", flush=True)
        if "until_assert" in k:
            # step both runs on until the left-edge assert fires in each; the
            # last step both finished is snapshotted again from a fresh run (a
            # run that asserted stops part-way through its step)
            trip = {}
            last = {8: t, 4: t}
            while len(trip) < 2 and t < k["until_assert"]:
                t += 1
                for prec, s in runs.items():
                    if prec in trip:
                        continue
                    try:
                        s.step(1)
                    except O.OracleError as e:
                        assert "prtls.rs:163-170" in str(e), str(e)
                        trip[prec] = t
                        continue
                    last[prec] = t
            for prec in (8, 4):
                out[f"{case}_assert_step_fp{prec * 8}"] = np.array(trip.get(prec, -1))
            # a common last good step for both precisions
            tl = min(last[8], last[4])
            steps.append(tl)
            for prec in (8, 4):
                s2 = oracle_run(case, prec, sp)
                s2.step(tl)
                snaps[prec].append(snapshot(s2, k))
            print(f"{case}: assert at step {trip} ({time.time() - tic:.0f} s)", flush=True)
        out[f"{case}_steps"] = np.array(steps)
        for name in ("ex", "ey", "bz"):
            out[f"{case}_{name}"] = np.stack([g[name] for g, _, _ in snaps[8]])
        for key in ("profile", "spec_ex", "spec_bz"):
            out[f"{case}_{key}"] = np.stack([m[key] for _, m, _ in snaps[8]])
        out[f"{case}_energy"] = np.stack([e for _, _, e in snaps[8]])
        # the FP32 oracle's errors against the FP64 one, per snapshot
        for key, val in errors_vs(out, case, [(m, g) for g, m, _ in snaps[4]]).items():
            out[f"{case}_fp32_oracle_{key}"] = val
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


def errors_vs(golden, case, states):
    """Errors of (metrics, interior fields) per snapshot against the FP64
    snapshots stored in `golden`."""
    k = CASES[case]
    errs = []
    for i, (m, g) in enumerate(states):
        snap = {key: golden[f"{case}_{key}"][i] for key in ("profile", "spec_ex", "spec_bz", "ex", "ey", "bz")}
        errs.append(snapshot_errors(snap, m, g, k["delta"]))
    return {key: np.array([e[key] for e in errs]) for key in errs[0]}


if __name__ == "__main__":
    main()
