"""FP32 against FP64 on the GPU: error of the FP32 path over a shock run.

    python -m spectral_rs_amd.precision [--size 512x128] [--steps 1000] [--every 50] [--json out.json]

Both runs start from the same state: the FP64 run is initialised on the
device (positions as src/prtls.rs:191-221, seeded momenta), and its particles
are rounded to FP32 for the FP32 run.  Every `every` steps the two are
compared:

  * fields: relative L2 error ||X32 - X64|| / ||X64|| of Ex, Ey, Bz, Jx, dens;
  * the shock density profile (dens averaged over y, as a function of x):
    relative L2 and max errors -- the quantity a shock study reads;
  * energy: total field + kinetic energy of each run relative to its start;
  * Gauss law: relative pre- and post-correction residuals of each run.

PIC trajectories are chaotic, so pointwise field differences grow until they
reach the particle-noise level; the y-averaged profile stays close much
longer.  FP64 on the GPU is the dprec path of the parity tests (bit-level
agreement with the FP64 restatement of the reference on the test fixture).
"""
import argparse
import json
import sys

import numpy as np

from spectral_rs_amd import capi

SHOCK = dict(delta=16, dens=32, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
SPECIES = [(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]  # ions then electrons


def _rel(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def _total(d):
    return d["energy_e"] + d["energy_b"] + sum(d["energy_kin"])


def run(size_x, size_y, steps, every, sort_interval=20, seed=0, out=sys.stdout):
    kw = dict(SHOCK, size_x=size_x, size_y=size_y)
    if size_x == 256:
        kw["delta"] = 8
    s64 = capi.GpuSim(prec=8, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT, **kw)
    s32 = capi.GpuSim(prec=4, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT, **kw)
    for k, (q, m, vth) in enumerate(SPECIES):
        sid = s64.init_species(q, m, vth, seed=seed)
        st = s64.species(sid)
        s32.add_species(q, m, st["ix"], st["iy"], st["dx"], st["dy"], st["px"], st["py"], st["pz"])
    for s in (s64, s32):
        s.prestep()
    e0 = {p: _total(s.diag()) for p, s in ((64, s64), (32, s32))}
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        s64.step(k)
        s32.step(k)
        done += k
        g64 = {n: s64.grid(n)[1:-1, 1:-1] for n in ("ex", "ey", "bz", "jx", "dens")}
        g32 = {n: s32.grid(n)[1:-1, 1:-1].astype(np.float64) for n in ("ex", "ey", "bz", "jx", "dens")}
        p64, p32 = g64["dens"].mean(axis=0), g32["dens"].mean(axis=0)
        d64, d32 = s64.diag(), s32.diag()
        row = {"step": done}
        row.update({f"l2_{n}": _rel(g32[n], g64[n]) for n in g64})
        row["profile_l2"] = _rel(p32, p64)
        row["profile_max"] = float(np.max(np.abs(p32 - p64)) / np.max(np.abs(p64)))
        row["energy_drift_f64"] = _total(d64) / e0[64] - 1.0
        row["energy_drift_f32"] = _total(d32) / e0[32] - 1.0
        row["gauss_pre_f64"], row["gauss_pre_f32"] = d64["gauss_pre"], d32["gauss_pre"]
        row["gauss_post_f64"], row["gauss_post_f32"] = d64["gauss_post_max"], d32["gauss_post_max"]
        rows.append(row)
        print(json.dumps(row), file=out, flush=True)
    s64.close()
    s32.close()
    return rows


def markdown(rows, size, steps):
    cols = ["step", "l2_ex", "l2_bz", "l2_jx", "l2_dens", "profile_l2", "profile_max", "energy_drift_f32",
            "energy_drift_f64", "gauss_pre_f32", "gauss_post_f32", "gauss_post_f64"]
    lines = [f"FP32 vs FP64 (GPU), shock {size}, {steps} steps", "",
             "| " + " | ".join(cols) + " |", "|" + "---|" * len(cols)]
    for r in rows:
        lines.append("| " + " | ".join(str(r[c]) if c == "step" else f"{r[c]:.3g}" for c in cols) + " |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", default="512x128")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--json", default=None)
    ap.add_argument("--markdown", default=None)
    a = ap.parse_args()
    sx, sy = (int(v) for v in a.size.lower().split("x"))
    rows = run(sx, sy, a.steps, a.every)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    if a.markdown:
        with open(a.markdown, "w") as f:
            f.write(markdown(rows, a.size, a.steps) + "\n")


if __name__ == "__main__":
    main()
