"""Precision study on the GPU: the FP32 product path ("mixed") and an
all-FP32 variant against the FP64 (dprec) path over a shock run.

    python -m spectral_rs_amd.precision [--size 2048x512] [--ppc 64] [--steps 5000] [--every 500] [--json out.json]

Variants:
  * mixed (the FP32 product path): positions as int32 cell + FP32 offset,
    FFT in FP32 with FP64-generated twiddles, energy/momentum sums in FP64;
  * all-FP32: the same kernels with the precision-study flags of sp_params:
    positions kept to the resolution of one absolute FP32 coordinate (after
    every position update the offset becomes fl(fl(ix + dx) - ix)), twiddles
    generated in FP32, diagnostic sums accumulated in FP32 (per thread and
    per wave);
  * all-FP64: the dprec path, the reference for the errors.

All three start from the same state: the FP64 run is initialised on the
device (positions as src/prtls.rs:191-221, seeded momenta) and its particles
are rounded to FP32 for the others.  Every `every` steps:

  * fields: relative L2 error ||X - X64|| / ||X64|| of Ex, Ey, Bz, Jx, dens;
  * the shock density profile (dens averaged over y): relative L2 and max
    errors -- the quantity a shock study reads;
  * energy: total field + kinetic energy of each run relative to its start;
  * Gauss law: relative pre- and post-correction residuals of each run.

PIC trajectories are chaotic, so pointwise field differences grow until they
reach the particle-noise level; the y-averaged profile stays close much
longer.
"""
import argparse
import json
import sys

import numpy as np

from spectral_rs_amd import capi

SHOCK = dict(n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
SPECIES = [(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]  # ions then electrons
ALL32 = capi.SP_STUDY_ABS_POS32 | capi.SP_STUDY_TWIDDLE32 | capi.SP_STUDY_SUM32
FIELDS = ("ex", "ey", "bz", "jx", "dens")


def _rel(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def _total(d):
    return d["energy_e"] + d["energy_b"] + sum(d["energy_kin"])


def run(size_x, size_y, ppc, steps, every, sort_interval=20, seed=0, out=sys.stdout):
    kw = dict(SHOCK, size_x=size_x, size_y=size_y, dens=ppc, delta=8 if size_x <= 256 else 16)
    sims = {"f64": capi.GpuSim(prec=8, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT, **kw),
            "mixed": capi.GpuSim(prec=4, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT, **kw),
            "all32": capi.GpuSim(prec=4, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT, study=ALL32,
                                 **kw)}
    for q, m, vth in SPECIES:
        sid = sims["f64"].init_species(q, m, vth, seed=seed)
        st = sims["f64"].species(sid)
        for v in ("mixed", "all32"):
            sims[v].add_species(q, m, st["ix"], st["iy"], st["dx"], st["dy"], st["px"], st["py"], st["pz"])
        del st
    for s in sims.values():
        s.prestep()
    e0 = {v: _total(s.diag()) for v, s in sims.items()}
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        for s in sims.values():
            s.step(k)
        done += k
        g = {v: {n: s.grid(n)[1:-1, 1:-1].astype(np.float64) for n in FIELDS} for v, s in sims.items()}
        d = {v: s.diag() for v, s in sims.items()}
        p64 = g["f64"]["dens"].mean(axis=0)
        row = {"step": done}
        for v in ("mixed", "all32"):
            row.update({f"l2_{n}_{v}": _rel(g[v][n], g["f64"][n]) for n in FIELDS})
            p = g[v]["dens"].mean(axis=0)
            row[f"profile_l2_{v}"] = _rel(p, p64)
            row[f"profile_max_{v}"] = float(np.max(np.abs(p - p64)) / np.max(np.abs(p64)))
        for v in sims:
            row[f"energy_drift_{v}"] = _total(d[v]) / e0[v] - 1.0
            row[f"gauss_pre_{v}"] = d[v]["gauss_pre"]
            row[f"gauss_post_max_{v}"] = d[v]["gauss_post_max"]
        rows.append(row)
        print(json.dumps(row), file=out, flush=True)
    for s in sims.values():
        s.close()
    return rows


def markdown(rows, size, ppc, steps):
    lines = [f"Precision study (GPU), shock {size}, {ppc} ppc per species, {steps} steps; errors against all-FP64", ""]
    cols = ["step"]
    for v in ("mixed", "all32"):
        cols += [f"l2_ex_{v}", f"l2_bz_{v}", f"l2_dens_{v}", f"profile_l2_{v}"]
    for v in ("mixed", "all32", "f64"):
        cols += [f"energy_drift_{v}", f"gauss_post_max_{v}"]
    lines += ["| " + " | ".join(cols) + " |", "|" + "---|" * len(cols)]
    for r in rows:
        lines.append("| " + " | ".join(str(r[c]) if c == "step" else f"{r[c]:.3g}" for c in cols) + " |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", default="2048x512")
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--json", default=None)
    ap.add_argument("--markdown", default=None)
    a = ap.parse_args()
    sx, sy = (int(v) for v in a.size.lower().split("x"))
    rows = run(sx, sy, a.ppc, a.steps, a.every)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    if a.markdown:
        with open(a.markdown, "w") as f:
            f.write(markdown(rows, a.size, a.ppc, a.steps) + "\n")


if __name__ == "__main__":
    main()
