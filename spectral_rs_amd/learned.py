"""Learned current denoiser: training data, training and the
time-to-solution experiment (host side of csrc/denoise.hip).

    python -m spectral_rs_amd.learned [--train-size 1024x256] [--eval-size 4096x1024] [--steps 1500] ...

1. Paired data from the method itself: 64-ppc shock runs (the plain method)
   with gamma0 drawn uniformly from [2, 20], each sampled every `every` steps.  At a sample, every 16th particle (by tag) of
   each species forms a 4-ppc copy of the same particle state; the copy moves
   and deposits (sp_move_and_deposit, ghost fold, the same binomial filter),
   giving the input, and the 64-ppc run's own filtered currents of that step
   are the target.  Both are divided by their ppc (the solve divides J and rho
   by dens, src/flds/mod.rs:80-84, :127) and by a per-channel scale.
2. Training on 64 x 64 patches (about 20k) with the BF16 MFMA kernels (Adam).
3. Held-out pairs: mean squared error of the binomial-filtered 4-ppc currents
   and of the network's output against the 64-ppc target.
4. Time to solution at the evaluation size: plain 64 ppc (the truth), plain
   4 ppc and learned 4 ppc, each run for `steps` steps; wall time of the step
   loop, the error of the y-averaged density profile (the shock's density
   jump) and of the x power spectra of Ex and Bz against the truth.
"""
import argparse
import json
import sys
import time

import numpy as np

from spectral_rs_amd import capi
from spectral_rs_amd.analysis import profile, profile_error, spectrum_error, xspectrum

SHOCK = dict(n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
SPECIES = [(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]  # ions then electrons
CH = ["jx", "jy", "jz", "rho"]
PATCH = 64


def delta_for(size_x):
    return 8 if size_x <= 256 else 16


def shock_kw(size_x, size_y, ppc):
    return dict(SHOCK, size_x=size_x, size_y=size_y, delta=delta_for(size_x), dens=ppc)


# the tile-sort schedule of every run here (the bench's): at least 3 and at
# most 40 steps apart, sooner when the particles spread out of the particle
# pass's windows
SORT_POLICY = (3, 40, 0.03)


def shock_sim(size_x, size_y, ppc, seed=0, sort_interval=20, gamma=10.0, sort_policy=SORT_POLICY):
    kw = shock_kw(size_x, size_y, ppc)
    kw["gamma_inj"] = gamma
    sim = capi.GpuSim(prec=4, sort_interval=sort_interval, left_edge=capi.SP_LEFT_REFLECT, **kw)
    if sort_policy:
        sim.set_sort_policy(*sort_policy)
    for q, m, vth in SPECIES:
        sim.init_species(q, m, vth, seed=seed)
    return sim


def currents(sim, ppc):
    """The 4 filtered current planes (interior) divided by ppc: [4][ny][nx]."""
    return np.stack([sim.grid(n)[1:-1, 1:-1] for n in CH]).astype(np.float32) / np.float32(ppc)


def pairs(size_x, size_y, ppc_hi=64, sub=16, t0=200, every=60, n=12, seed=0, gamma=10.0, log=sys.stderr):
    """Current planes [n][4][ny][nx] from one 64-ppc run: the 4-ppc copy's
    raw deposit (ghost-folded, unfiltered), the same after the binomial
    filter, and the 64-ppc filtered target."""
    hi = shock_sim(size_x, size_y, ppc_hi, seed, gamma=gamma)
    hi.prestep()
    hi.step(t0)
    ppc_lo = ppc_hi // sub
    X, XF, T = [], [], []
    for k in range(n):
        hi.sync()
        kw = shock_kw(size_x, size_y, ppc_lo)
        kw["gamma_inj"] = gamma
        lo = capi.GpuSim(prec=4, sort_interval=0, left_edge=capi.SP_LEFT_REFLECT, **kw)
        for sid, (q, m, _) in enumerate(SPECIES):
            st = hi.species(sid)
            sel = slice(k % sub, None, sub)  # every sub-th particle of the state
            lo.add_species(q, m, st["ix"][sel], st["iy"][sel], st["dx"][sel], st["dy"][sel], st["px"][sel],
                           st["py"][sel], st["pz"][sel])
        for sid in range(len(SPECIES)):
            lo.move_and_deposit(sid)
        lo.fold_all()
        X.append(currents(lo, ppc_lo))
        for name in CH:
            lo.filter(name, SHOCK["n_pass"])
        XF.append(currents(lo, ppc_lo))
        lo.close()
        hi.step(1)  # the same move and deposit at 64 ppc, then its field solve (filter included)
        hi.sync()
        T.append(currents(hi, ppc_hi))
        hi.step(every - 1)
        print(f"pairs: sample {k + 1}/{n} at step {t0 + k * every}", file=log, flush=True)
    hi.close()
    return np.stack(X), np.stack(XF), np.stack(T)


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


def layer_shapes():
    """(cout, cin) of each convolution: 4 -> 32 -> 32 -> 32 -> 32 -> 4."""
    return [(capi.DN_CH if l < capi.DN_LAYERS - 1 else 4, 4 if l == 0 else capi.DN_CH)
            for l in range(capi.DN_LAYERS)]


def init_params(seed, scales):
    """He-style initial weights, zero biases, small last layer; channel scales."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (co, ci) in enumerate(layer_shapes()):
        w = rng.standard_normal((co, ci * 9)) * np.sqrt(2.0 / (ci * 9))
        if l == capi.DN_LAYERS - 1:
            w *= 0.1
        parts += [w.ravel(), np.zeros(co)]
    parts.append(np.asarray(scales, np.float64))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == capi.DN_NPARAMS
    return p


def train(X, T, iters=2000, lr=2e-3, batch=16, seed=0, per_sample=64, log=sys.stderr):
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
    xs, ts = patches(Xn, Tn, rng, per_sample=per_sample)
    dn = capi.Denoiser(batch, PATCH, PATCH)
    dn.set_params(init_params(seed, scales))
    dn.set_loss_weights((1.0 / (4.0 * noise)).astype(np.float32))
    # step decay of the learning rate: half the steps at lr, then a quarter at
    # lr/2, lr/4 for an eighth and lr/10 for the rest (Adam state carries over)
    plan = [(iters // 2, lr), (iters // 4, lr / 2), (iters // 8, lr / 4)]
    plan.append((iters - sum(n for n, _ in plan), lr / 10))
    hist = []
    first = True
    for k, (n, r) in enumerate(plan):
        for j in range(0, n, 500):  # in chunks, with a progress line each
            m = min(500, n - j)
            hist.append(dn.train(xs if first else None, ts if first else None, m, r, seed=seed + 1000 * k + j))
            first = False
            print(f"train: lr {r:.3g}, {sum(h.size for h in hist)}/{iters} steps, loss {hist[-1][-50:].mean():.4g}",
                  file=log, flush=True)
    hist = np.concatenate(hist)
    p = dn.params()
    dn.close()
    print(f"train: {iters} Adam steps on {xs.shape[0]} patches, loss {hist[:20].mean():.4g} -> "
          f"{hist[-50:].mean():.4g}", file=log, flush=True)
    return p, hist, scales


def evaluate(p, X, XF, T):
    """MSE (per channel, normalised by the target's mean square) of the
    binomial-filtered 4-ppc currents XF and of the network's output on its
    input X (the raw or the filtered deposit), against the 64-ppc target."""
    scales = p[-4:].astype(np.float64)
    n, _, ny, nx = X.shape
    dn = capi.Denoiser(1, ny, nx)
    dn.set_params(p)
    base, learned = np.zeros(4), np.zeros(4)
    for i in range(n):
        xn = (X[i] / scales[:, None, None]).astype(np.float32)[None]
        y = dn.forward(xn, mode=1)[0] * scales[:, None, None]
        base += np.mean((XF[i] - T[i]) ** 2, axis=(1, 2))
        learned += np.mean((y - T[i]) ** 2, axis=(1, 2))
    dn.close()
    ms = np.mean(T.astype(np.float64) ** 2, axis=(0, 2, 3))
    return {"channels": CH, "binomial_rel_mse": list(base / n / ms), "learned_rel_mse": list(learned / n / ms)}


def run_case(size_x, size_y, ppc, steps, params=None, seed=0, max_rel=1.0, mask=0xF, raw=True, max_hp=0.0, backoff=0):
    sim = shock_sim(size_x, size_y, ppc, seed)
    if params is not None:
        # the network may remove most of a channel that is almost pure noise
        # (Jz), so the correction is allowed up to the size of the input; the
        # checks are on this solve's data (no gate on the previous Gauss residual)
        sim.set_denoiser(params, max_rel=max_rel, max_gauss=0.0)
        sim.set_denoiser_input(raw, max_hp)
        sim.set_denoiser_channels(mask)
        sim.set_denoiser_backoff(backoff)
    sim.prestep()
    sim.step(2)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    el = time.perf_counter() - t0
    prof = profile(sim, ppc)
    spec = {n: xspectrum(sim, n) for n in ("ex", "bz")}
    stats = (*sim.denoiser_stats(), sim.denoiser_skipped()) if params is not None else None
    sim.close()
    return el, prof, stats, spec


def heldout(p, size_x, size_y, gammas, seed0, n, t0, every, raw, log=sys.stderr):
    """Network against the binomial filter on pairs from runs the training
    never saw: one run per gamma0, seeds seed0, seed0 + 1, ..."""
    out = []
    for k, g in enumerate(gammas):
        Xr, Xf, T = pairs(size_x, size_y, n=n, t0=t0, every=every, seed=seed0 + k, gamma=g, log=log)
        ev = evaluate(p, Xr if raw else Xf, Xf, T)
        ev.update(gamma=g, seed=seed0 + k)
        print("held-out:", json.dumps(ev), file=log, flush=True)
        out.append(ev)
    return out


def equal_error(res, learned, metrics=("profile_rel_err", "spectrum_rel_err_ex", "spectrum_rel_err_bz")):
    """For each metric: the plain particle count whose error equals the
    learned run's (log-log interpolation between the plain runs; a bound when
    outside their range) and the time to solution of that plain run
    (interpolated the same way) over the learned run's."""
    plain = sorted((int(k.split("_")[1][:-3]), v) for k, v in res.items()
                   if k.startswith("plain_") and k != "plain_64ppc" and metrics[0] in v)
    if len(plain) < 2:
        return {}
    ppc = np.log([c for c, _ in plain])
    tl = np.log([v["s_per_run"] for _, v in plain])
    out = {}
    for m in metrics:
        e = np.log([max(v[m], 1e-300) for _, v in plain])
        el = np.log(max(res[learned][m], 1e-300))
        # errors fall with ppc: interpolate ppc as a function of error
        order = np.argsort(e)
        pe = float(np.interp(el, e[order], ppc[order]))
        if el < e.min():
            note = "learned error below every plain run's: plain ppc and time are lower bounds"
        elif el > e.max():
            note = "learned error above every plain run's: plain ppc and time are upper bounds"
        else:
            note = "interpolated"
        t = float(np.exp(np.interp(pe, ppc, tl)))
        out[m] = {"plain_ppc_at_equal_error": float(np.exp(pe)), "plain_s_per_run": t,
                  "learned_s_per_run": res[learned]["s_per_run"], "speedup": t / res[learned]["s_per_run"],
                  "note": note}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-size", default="1024x256")
    ap.add_argument("--eval-size", default="4096x1024")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--train-runs", type=int, default=12, help="training runs, gamma0 uniform in [2, 20]")
    ap.add_argument("--train-seed", type=int, default=0, help="seed of the first training run and of the gamma0 draw")
    ap.add_argument("--samples", type=int, default=12, help="pairs per training run")
    ap.add_argument("--patches", type=int, default=140, help="64x64 patches per pair")
    ap.add_argument("--input", choices=["raw", "filtered"], default="raw",
                    help="the network's input: the raw 4-ppc deposit (it stands in for the binomial filter) "
                         "or the filtered one")
    ap.add_argument("--max-hp", type=float, default=2.0,
                    help="spectrum check: high-pass energy of the output at most this times the filter's")
    ap.add_argument("--heldout-seed", type=int, default=100)
    ap.add_argument("--heldout-gammas", default="4.0,12.9,18.0", help="gamma0 of the held-out runs")
    ap.add_argument("--t0", type=int, default=100)
    ap.add_argument("--every", type=int, default=75)
    ap.add_argument("--iters", type=int, default=12000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-seeds", default="0,1", help="runs per case; profiles are averaged over them")
    ap.add_argument("--cases", default="plain_64ppc,plain_4ppc,plain_8ppc,plain_16ppc,learned",
                    help="cases of the time-to-solution study (plain_<n>ppc, learned)")
    ap.add_argument("--dn-masks", default="7",
                    help="channel masks of the learned cases (bit c: Jx Jy Jz rho written back)")
    ap.add_argument("--dn-backoffs", default="0,16",
                    help="backoffs of the learned cases: solves that keep the filter's result without inference after "
                         "a failed check")
    ap.add_argument("--out", default=None, help="JSON report")
    ap.add_argument("--params-out", default=None, help=".npy of the trained parameters")
    ap.add_argument("--params-in", default=None, help="skip training: evaluate these parameters (.npy)")
    a = ap.parse_args()
    tx, ty = (int(v) for v in a.train_size.split("x"))
    ex, ey = (int(v) for v in a.eval_size.split("x"))
    raw = a.input == "raw"
    hg = [float(v) for v in a.heldout_gammas.split(",")] if a.heldout_gammas else []
    if a.params_in:
        p = np.load(a.params_in)
        rep = {"params": a.params_in, "input": a.input}
    else:
        grng = np.random.default_rng(a.train_seed)
        gammas = [float(g) for g in grng.uniform(2.0, 20.0, a.train_runs)]
        seeds = [a.train_seed + k for k in range(a.train_runs)]
        XT = [pairs(tx, ty, n=a.samples, t0=a.t0, every=a.every, seed=sd, gamma=g) for sd, g in zip(seeds, gammas)]
        X = np.concatenate([(x if raw else xf) for x, xf, _ in XT])
        XF = np.concatenate([xf for _, xf, _ in XT])
        T = np.concatenate([t for _, _, t in XT])
        del XT
        p, hist, scales = train(X, T, iters=a.iters, lr=a.lr, per_sample=a.patches)
        npatch = X.shape[0] * a.patches
        # the same full-frame measure on frames the network was trained on
        pick = np.linspace(0, X.shape[0] - 1, min(8, X.shape[0])).astype(int)
        ev_train = evaluate(p, X[pick], XF[pick], T[pick])
        print("training frames:", json.dumps(ev_train), file=sys.stderr, flush=True)
        del X, XF, T
        rep = {"train_size": a.train_size, "input": a.input, "seeds": seeds, "gammas": gammas,
               "samples_per_run": a.samples, "patches": npatch, "t0": a.t0, "every": a.every, "iters": a.iters,
               "lr": a.lr, "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-50:].mean()),
               "training_frames": ev_train}
    if hg:
        rep["heldout"] = heldout(p, tx, ty, hg, a.heldout_seed, max(4, a.samples // 2), a.t0, a.every * 2, raw)
    d = delta_for(ex)
    eseeds = [int(v) for v in a.eval_seeds.split(",")]
    masks = [int(v) for v in a.dn_masks.split(",")]
    backoffs = [int(v) for v in a.dn_backoffs.split(",")]
    want = [c for c in a.cases.split(",") if c]
    cases = {}
    for c in want:
        if c.startswith("plain_"):
            cases[c] = (int(c[len("plain_"):-3]), None, 0, 0)
    if "learned" in want:
        for mk in masks:
            for bo in backoffs:
                cases[f"learned_4ppc_mask{mk}" + (f"_backoff{bo}" if bo else "")] = (4, p, mk, bo)
    res = {}
    for name, (ppc, prm, mk, bo) in cases.items():
        runs = [run_case(ex, ey, ppc, a.steps, params=prm, seed=sd, mask=mk, raw=raw, max_hp=a.max_hp, backoff=bo)
                for sd in eseeds]
        res[name] = dict(t=float(np.mean([r[0] for r in runs])), prof=np.mean([r[1] for r in runs], axis=0),
                         stats=[r[2] for r in runs],
                         spec={n: np.mean([r[3][n] for r in runs], axis=0) for n in ("ex", "bz")})
        print(f"time_to_solution: {name} {res[name]['t']:.3f} s per run", file=sys.stderr, flush=True)
    truth = res.get("plain_64ppc")
    tts = {"eval_size": a.eval_size, "steps": a.steps, "seeds": eseeds, "max_hp": a.max_hp, "cases": {}}
    for case in cases:
        r = res[case]
        e = {"s_per_run": r["t"]}
        if truth is not None and case != "plain_64ppc":
            e["speedup_vs_plain_64ppc"] = truth["t"] / r["t"]
            e["profile_rel_err"] = profile_error(r["prof"], truth["prof"], d)
            for n in ("ex", "bz"):
                e[f"spectrum_rel_err_{n}"] = spectrum_error(r["spec"][n], truth["spec"][n])
        if cases[case][1] is not None:
            e["applied_fallbacks"] = r["stats"]
        tts["cases"][case] = e
    for case in cases:
        if case.startswith("learned") and truth is not None:
            tts["cases"][case]["equal_error_vs_plain"] = equal_error(tts["cases"], case)
    rep["time_to_solution"] = tts
    print(json.dumps(rep), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)
    if a.params_out:
        np.save(a.params_out, p)


if __name__ == "__main__":
    main()
