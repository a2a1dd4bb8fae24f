"""Seeded synthetic shock inputs, restated with numpy in Float arithmetic.

* positions: Prtl::initialize_positions (src/prtls.rs:191-221): every cell of
  the slab delta <= j < size_x - delta holds dens particles on the cell
  diagonal, dx = dy = 1/(2 dens) + k/dens - 1/2, in ghost-offset cells;
* momenta: the recipe of Prtl::initialize_velocities (src/prtls.rs:222-250)
  from numpy's seeded generator instead of the unseeded thread_rng.

The GPU has its own counter-based initialiser (sp_species_init); these arrays
are what host-side callers, tests and the oracle runs upload identically to
every side.
"""
import numpy as np


def positions(size_x, size_y, delta, dens, ft=np.float32):
    """src/prtls.rs:191-221, restated with numpy in Float arithmetic."""
    ft = np.dtype(ft).type
    i = np.repeat(np.arange(size_y), (size_x - 2 * delta) * dens)
    j = np.tile(np.repeat(np.arange(delta, size_x - delta), dens), size_y)
    k = np.tile(np.arange(dens), (size_x - 2 * delta) * size_y)
    r1 = ft(1.0) / (ft(2.0) * ft(dens))
    r1 = (r1 + k.astype(ft) / ft(dens)).astype(ft)
    d = (r1 - ft(0.5)).astype(ft)
    return (j + 1).astype(np.int64), (i + 1).astype(np.int64), d.copy(), d.copy()


def momenta(n, c, gamma_inj, vth, seed, ft=np.float32):
    """Seeded stand-in for src/prtls.rs:222-250: Gaussian momenta of width
    vth*c boosted by the Zenitani flip, in Float arithmetic.  The reference
    draws from an unseeded thread_rng; only the recipe is reproduced."""
    ft = np.dtype(ft).type
    rng = np.random.default_rng(seed)
    c = ft(c)
    g = ft(gamma_inj)
    csqinv = ft(1.0) / (c * c)
    beta_inj = np.sqrt(ft(1.0) - ft(1.0) / (g * g)).astype(ft)
    s = ft(vth) * c
    px = (rng.standard_normal(n).astype(ft) * s).astype(ft)
    py = (rng.standard_normal(n).astype(ft) * s).astype(ft)
    pz = (rng.standard_normal(n).astype(ft) * s).astype(ft)
    psa = np.sqrt(ft(1.0) + (px * px + py * py + pz * pz) * csqinv).astype(ft)
    ux = (px / c).astype(ft)
    r = rng.random(n).astype(ft)
    flip = (-beta_inj * ux) > (r * psa)
    ux = np.where(flip, -ux, ux).astype(ft)
    px = (g * (ux + beta_inj * psa)).astype(ft)
    px = (px * c).astype(ft)
    psa = np.sqrt(ft(1.0) + (px * px + py * py + pz * pz) * csqinv).astype(ft)
    return px, py, pz, psa


def shock_species(size_x, size_y, delta, dens, c, gamma_inj, vth, seed, ft=np.float32):
    ix, iy, dx, dy = positions(size_x, size_y, delta, dens, ft)
    px, py, pz, psa = momenta(ix.size, c, gamma_inj, vth, seed, ft)
    return dict(ix=ix, iy=iy, dx=dx, dy=dy, px=px, py=py, pz=pz, psa=psa)
