"""Learned current denoiser: training data, training and the
time-to-solution experiment (host side of csrc/denoise.hip).

    python -m spectral_rs_amd.learned [--train-size 512x128] [--eval-size 2048x512] [--steps 1500] ...

1. Paired data from the method itself: a 64-ppc shock run (the plain method),
   sampled every `every` steps.  At a sample, every 16th particle (by tag) of
   each species forms a 4-ppc copy of the same particle state; the copy moves
   and deposits (sp_move_and_deposit, ghost fold, the same binomial filter),
   giving the input, and the 64-ppc run's own filtered currents of that step
   are the target.  Both are divided by their ppc (the solve divides J and rho
   by dens, src/flds/mod.rs:80-84, :127) and by a per-channel scale.
2. Training on 64 x 64 patches with the BF16 MFMA kernels (Adam).
3. Held-out pairs: mean squared error of the binomial-filtered 4-ppc currents
   and of the network's output against the 64-ppc target.
4. Time to solution at the evaluation size: plain 64 ppc (the truth), plain
   4 ppc and learned 4 ppc, each run for `steps` steps; wall time of the step
   loop and the error of the y-averaged density profile against the truth.
"""
import argparse
import json
import sys
import time

import numpy as np

from spectral_rs_amd import capi

SHOCK = dict(n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
SPECIES = [(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]  # ions then electrons
CH = ["jx", "jy", "jz", "rho"]
PATCH = 64


def delta_for(size_x):
    return 8 if size_x <= 256 else 16


def shock_kw(size_x, size_y, ppc):
    return dict(SHOCK, size_x=size_x, size_y=size_y, delta=delta_for(size_x), dens=ppc)


def shock_sim(size_x, size_y, ppc, seed=0, sort_interval=20):
    sim = capi.GpuSim(prec=4, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT,
                      **shock_kw(size_x, size_y, ppc))
    for q, m, vth in SPECIES:
        sim.init_species(q, m, vth, seed=seed)
    return sim


def currents(sim, ppc):
    """The 4 filtered current planes (interior) divided by ppc: [4][ny][nx]."""
    return np.stack([sim.grid(n)[1:-1, 1:-1] for n in CH]).astype(np.float32) / np.float32(ppc)


def pairs(size_x, size_y, ppc_hi=64, sub=16, t0=200, every=60, n=12, seed=0, log=sys.stderr):
    """Input / target current planes [n][4][ny][nx] from one 64-ppc run."""
    hi = shock_sim(size_x, size_y, ppc_hi, seed)
    hi.prestep()
    hi.step(t0)
    ppc_lo = ppc_hi // sub
    X, T = [], []
    for k in range(n):
        hi.sync()
        lo = capi.GpuSim(prec=4, sort_interval=0, left_edge=capi.SP_LEFT_REFLECT,
                         **shock_kw(size_x, size_y, ppc_lo))
        for sid, (q, m, _) in enumerate(SPECIES):
            st = hi.species(sid)
            sel = slice(k % sub, None, sub)  # every sub-th particle of the state
            lo.add_species(q, m, st["ix"][sel], st["iy"][sel], st["dx"][sel], st["dy"][sel], st["px"][sel],
                           st["py"][sel], st["pz"][sel])
        for sid in range(len(SPECIES)):
            lo.move_and_deposit(sid)
        lo.fold_all()
        for name in CH:
            lo.filter(name, SHOCK["n_pass"])
        X.append(currents(lo, ppc_lo))
        lo.close()
        hi.step(1)  # the same move and deposit at 64 ppc, then its field solve (filter included)
        hi.sync()
        T.append(currents(hi, ppc_hi))
        hi.step(every - 1)
        print(f"pairs: sample {k + 1}/{n} at step {t0 + k * every}", file=log, flush=True)
    hi.close()
    return np.stack(X), np.stack(T)


def patches(X, T, rng, per_sample=64, size=PATCH):
    """Random size x size patches of the pairs."""
    n, _, ny, nx = X.shape
    xs, ts = [], []
    for i in range(n):
        for _ in range(per_sample):
            y0 = rng.integers(0, ny - size + 1)
            x0 = rng.integers(0, nx - size + 1)
            xs.append(X[i, :, y0:y0 + size, x0:x0 + size])
            ts.append(T[i, :, y0:y0 + size, x0:x0 + size])
    return np.ascontiguousarray(np.stack(xs)), np.ascontiguousarray(np.stack(ts))


def init_params(seed, scales):
    """He-style initial weights, zero biases, small last layer; channel scales."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in [(16, 4 * 9), (16,), (16, 16 * 9), (16,), (4, 16 * 9), (4,)]:
        if len(shape) == 2:
            w = rng.standard_normal(shape) * np.sqrt(2.0 / shape[1])
            if shape[0] == 4:
                w *= 0.1
            parts.append(w.ravel())
        else:
            parts.append(np.zeros(shape))
    parts.append(np.asarray(scales, np.float64))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == capi.DN_NPARAMS
    return p


def train(X, T, iters=2000, lr=2e-3, batch=16, seed=0, log=sys.stderr):
    """Trains on patches of the pairs; returns (params, loss history, scales).
    Each channel is scaled by the RMS of the input (values of order 1 in
    BF16), and its loss is weighted by 1 / (its scaled 4-ppc noise power), so
    that every channel's noise counts equally (Jz is almost pure noise)."""
    X64 = X.astype(np.float64)
    scales = np.sqrt(np.mean(X64 ** 2, axis=(0, 2, 3))) + 1e-12
    noise = np.mean((X64 - T) ** 2, axis=(0, 2, 3)) / scales ** 2
    Xn = X / scales[None, :, None, None].astype(np.float32)
    Tn = T / scales[None, :, None, None].astype(np.float32)
    rng = np.random.default_rng(seed)
    xs, ts = patches(Xn, Tn, rng)
    dn = capi.Denoiser(batch, PATCH, PATCH)
    dn.set_params(init_params(seed, scales))
    dn.set_loss_weights((1.0 / (4.0 * noise)).astype(np.float32))
    hist = dn.train(xs, ts, iters, lr, seed=seed)
    p = dn.params()
    dn.close()
    print(f"train: {iters} Adam steps on {xs.shape[0]} patches, loss {hist[:20].mean():.4g} -> "
          f"{hist[-50:].mean():.4g}", file=log, flush=True)
    return p, hist, scales


def evaluate(p, X, T):
    """MSE (per channel, normalised by the target's mean square) of the binomial
    4-ppc currents and of the network's output, against the 64-ppc target."""
    scales = p[-4:].astype(np.float64)
    n, _, ny, nx = X.shape
    dn = capi.Denoiser(1, ny, nx)
    dn.set_params(p)
    base, learned = np.zeros(4), np.zeros(4)
    for i in range(n):
        xn = (X[i] / scales[:, None, None]).astype(np.float32)[None]
        y = dn.forward(xn, mode=1)[0] * scales[:, None, None]
        base += np.mean((X[i] - T[i]) ** 2, axis=(1, 2))
        learned += np.mean((y - T[i]) ** 2, axis=(1, 2))
    dn.close()
    ms = np.mean(T.astype(np.float64) ** 2, axis=(0, 2, 3))
    return {"channels": CH, "binomial_rel_mse": list(base / n / ms), "learned_rel_mse": list(learned / n / ms)}


def profile(sim, ppc):
    """Density averaged over y, relative to the upstream ppc."""
    return sim.grid("dens")[1:-1, 1:-1].astype(np.float64).mean(axis=0) / ppc


def run_case(size_x, size_y, ppc, steps, params=None, seed=0, max_rel=1.0):
    sim = shock_sim(size_x, size_y, ppc, seed)
    if params is not None:
        # the network may remove most of a channel that is almost pure noise
        # (Jz), so the correction is allowed up to the size of the input
        sim.set_denoiser(params, max_rel=max_rel)
    sim.prestep()
    sim.step(2)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    el = time.perf_counter() - t0
    prof = profile(sim, ppc)
    stats = sim.denoiser_stats() if params is not None else None
    sim.close()
    return el, prof, stats


def profile_error(prof, truth, delta):
    a, b = prof[delta:-delta], truth[delta:-delta]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-size", default="512x128")
    ap.add_argument("--eval-size", default="2048x512")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--samples", type=int, default=16, help="pairs per training run")
    ap.add_argument("--seeds", default="0,1", help="seeds of the training runs")
    ap.add_argument("--heldout-seed", type=int, default=5)
    ap.add_argument("--t0", type=int, default=100)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-seeds", default="0,1,2", help="runs per case; profiles are averaged over them")
    ap.add_argument("--out", default=None, help="JSON report")
    ap.add_argument("--params-out", default=None, help=".npy of the trained parameters")
    ap.add_argument("--params-in", default=None, help="skip training: evaluate these parameters (.npy)")
    a = ap.parse_args()
    tx, ty = (int(v) for v in a.train_size.split("x"))
    ex, ey = (int(v) for v in a.eval_size.split("x"))
    if a.params_in:
        p = np.load(a.params_in)
        rep = {"params": a.params_in}
    else:
        seeds = [int(v) for v in a.seeds.split(",")]
        XT = [pairs(tx, ty, n=a.samples, t0=a.t0, every=a.every, seed=sd) for sd in seeds]
        X, T = np.concatenate([x for x, _ in XT]), np.concatenate([t for _, t in XT])
        p, hist, scales = train(X, T, iters=a.iters, lr=a.lr)
        # held out: a run with another seed, sampled over the same times
        Xh, Th = pairs(tx, ty, n=max(4, a.samples // 4), t0=a.t0, every=a.every * 4, seed=a.heldout_seed)
        ev = evaluate(p, Xh, Th)
        print("held-out:", json.dumps(ev), file=sys.stderr, flush=True)
        rep = {"train_size": a.train_size, "seeds": seeds, "samples_per_run": a.samples, "t0": a.t0,
               "every": a.every, "iters": a.iters, "loss_first": float(hist[:20].mean()),
               "loss_last": float(hist[-50:].mean()), "heldout_seed": a.heldout_seed, "heldout": ev}
    d = delta_for(ex)
    eseeds = [int(v) for v in a.eval_seeds.split(",")]
    cases = {"plain_64ppc": (64, None), "plain_4ppc": (4, None), "learned_4ppc": (4, p)}
    res = {}
    for name, (ppc, prm) in cases.items():
        runs = [run_case(ex, ey, ppc, a.steps, params=prm, seed=sd) for sd in eseeds]
        res[name] = (float(np.mean([r[0] for r in runs])), np.mean([r[1] for r in runs], axis=0),
                     [r[2] for r in runs])
        print(f"time_to_solution: {name} {res[name][0]:.3f} s per run", file=sys.stderr, flush=True)
    truth = res["plain_64ppc"][1]
    rep["time_to_solution"] = {
        "eval_size": a.eval_size, "steps": a.steps, "seeds": eseeds,
        "plain_64ppc_s": res["plain_64ppc"][0], "plain_4ppc_s": res["plain_4ppc"][0],
        "learned_4ppc_s": res["learned_4ppc"][0],
        "speedup_learned_vs_plain64": res["plain_64ppc"][0] / res["learned_4ppc"][0],
        "profile_rel_err_plain_4ppc": profile_error(res["plain_4ppc"][1], truth, d),
        "profile_rel_err_learned_4ppc": profile_error(res["learned_4ppc"][1], truth, d),
        "denoiser_applied_fallbacks": res["learned_4ppc"][2]}
    print(json.dumps(rep), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)
    if a.params_out:
        np.save(a.params_out, p)


if __name__ == "__main__":
    main()
