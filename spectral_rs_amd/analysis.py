"""Shock diagnostics used to compare runs (no reference counterpart beyond
plotter.py:9-19, which plots the y-averaged density of the output files).

Each takes anything with a ``grid(name)`` getter returning the padded
(Ny+2) x (Nx+2) array, so GPU contexts and saved states read the same way.

* density profile: the dens grid averaged over y, in units of the upstream
  particles per cell of one species -- the shock's density jump;
* x power spectra of a field, averaged over rows;
* relative L2 errors of these against a reference run.
"""
import numpy as np


def profile(sim, ppc):
    """Density averaged over y, relative to the upstream ppc (both species
    deposit into dens, so the upstream value is 2)."""
    return sim.grid("dens")[1:-1, 1:-1].astype(np.float64).mean(axis=0) / ppc


def xspectrum_of(field):
    """|FFT_x|^2 of an interior field, averaged over rows, for kx = 0..Nx/2."""
    return np.mean(np.abs(np.fft.rfft(np.asarray(field, np.float64), axis=1)) ** 2, axis=0)


def xspectrum(sim, name):
    return xspectrum_of(sim.grid(name)[1:-1, 1:-1])


def profile_error(prof, truth, delta):
    """Relative L2 error of a density profile between the slab's ends."""
    a, b = prof[delta:-delta], truth[delta:-delta]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def spectrum_error(spec, truth):
    """Relative L2 error of a power spectrum over kx = 1 .. Nx/4 (the modes the
    binomial filter leaves)."""
    k = slice(1, (truth.size - 1) // 2 + 1)
    return float(np.linalg.norm(spec[k] - truth[k]) / np.linalg.norm(truth[k]))


def field_error(f, truth):
    """Relative L2 error of an interior field."""
    a, b = np.asarray(f, np.float64), np.asarray(truth, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def shock_metrics(sim, ppc):
    """Profile and Ex, Bz spectra of a state."""
    return {"profile": profile(sim, ppc), "spec_ex": xspectrum(sim, "ex"), "spec_bz": xspectrum(sim, "bz")}


def snapshot_errors(snap, metrics, fields, delta):
    """Errors of one state (its shock_metrics and interior Ex, Ey, Bz) against
    a reference snapshot holding the same keys."""
    e = {"profile": profile_error(metrics["profile"], snap["profile"], delta),
         "spec_ex": spectrum_error(metrics["spec_ex"], snap["spec_ex"]),
         "spec_bz": spectrum_error(metrics["spec_bz"], snap["spec_bz"])}
    for n in ("ex", "ey", "bz"):
        e[n] = field_error(fields[n], snap[n])
    return e


def load_snapshots(path, case):
    """The snapshots of one case of a make_shock_snapshots file, as a list of
    dicts (one per step) plus the case's scalars."""
    with np.load(path, allow_pickle=False) as z:
        d = {k[len(case) + 1:]: z[k] for k in z.files if k.startswith(case + "_")}
    steps = [int(s) for s in d["steps"]]
    snaps = [{k: d[k][i] for k in ("profile", "spec_ex", "spec_bz", "ex", "ey", "bz", "energy")}
             for i in range(len(steps))]
    return steps, snaps, d


# ---- long-horizon comparisons against stored snapshots of another run ----
# The shock workloads of the stored FP64 snapshots (reference units: c = 10,
# dt = 0.045 -> 0.45 cells/step, 10-cell skin depth; gamma0 = 10; 4 filter
# passes; ions q = +1, m = 16, vth 0.0025, then electrons q = -1, m = 1,
# vth 0.01, inputs from spectral_rs_amd.inputs.shock_species with seeds 0, 1).
SHOCK_PHYS = dict(n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
SHOCK_SPECIES = [(1.0, 16.0, 0.0025, 0), (-1.0, 1.0, 0.01, 1)]
HORIZON_CASES = {
    # the 512 x 128 slab of the 4096 x 1024 shock (32 ppc, 3,932,160 particles)
    "c2_crop": dict(size_x=512, size_y=128, delta=16, dens=32, steps=[100, 500, 1000], left_edge=1),
    # the 256 x 64 shock at 16 ppc (491,520 particles), to the left-edge assert
    "c1": dict(size_x=256, size_y=64, delta=8, dens=16, steps=[100, 300], left_edge=0, until_assert=1000),
}


def shock_inputs(case, ft=np.float32):
    """[(charge, mass, arrays)] of a HORIZON_CASES workload."""
    from spectral_rs_amd.inputs import shock_species
    k = HORIZON_CASES[case]
    return [(q, m, shock_species(k["size_x"], k["size_y"], k["delta"], k["dens"], SHOCK_PHYS["c"],
                                 SHOCK_PHYS["gamma_inj"], vth, seed=sd, ft=ft)) for q, m, vth, sd in SHOCK_SPECIES]


def inputs_digest(sp):
    import hashlib
    h = hashlib.sha256()
    for _, _, a in sp:
        for key in ("ix", "iy", "dx", "dy", "px", "py", "pz"):
            h.update(np.ascontiguousarray(a[key]).tobytes())
    return h.hexdigest()


def horizon_errors(sim, case, path):
    """Steps `sim` (species added, pre-step done, at step 0) through the
    snapshot steps of `case` stored in `path` and returns, per step, its
    errors against them (snapshot_errors keys) and the stored FP32-oracle
    errors of the same step."""
    k = HORIZON_CASES[case]
    steps, snaps, d = load_snapshots(path, case)
    out, t = [], 0
    for i, st in enumerate(steps):
        sim.step(st - t)
        t = st
        m = shock_metrics(sim, k["dens"])
        f = {n: sim.grid(n)[1:-1, 1:-1] for n in ("ex", "ey", "bz")}
        e = snapshot_errors(snaps[i], m, f, k["delta"])
        ref = {key: float(d[f"fp32_oracle_{key}"][i]) for key in e}
        out.append({"step": st, "errors": e, "oracle_fp32_errors": ref})
    return out
