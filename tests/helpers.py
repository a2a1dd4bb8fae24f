"""Shared test inputs: the reference's test fixture and seeded shock set-ups."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ref_vectors():
    """Golden vectors from the reference's unit tests, keyed by (test, label)."""
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        d = json.load(f)
    return {(v["test"], v["label"]): np.array(v["values"], dtype=np.float64) for v in d["vectors"]}


# src/lib.rs:219-244 (build_test_sim)
TEST_SIM = dict(size_x=24, size_y=12, delta=5, dens=2, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0)
# src/lib.rs:26-32
E_TOL = {4: 1e-4, 8: 1e-13}


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
