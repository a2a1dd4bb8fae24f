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
