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


# Cold Langmuir oscillation: electrons with px = u0 c sin(2 pi x / nx) on
# heavy ions at the same places (zero net charge at t=0), periodic in x and y.
LANGMUIR = dict(size_x=32, size_y=8, delta=0, dens=4, n_pass=0, dt=0.1, c=3.0, gamma_inj=5.0, periodic_x=1)
LANGMUIR_ION_MASS = 1e4


def langmuir_species(u0=0.05, ft=np.float64):
    """[(charge, mass, arrays)] for the ions then the electrons of LANGMUIR."""
    L = LANGMUIR
    ix, iy, dx, dy = positions(L["size_x"], L["size_y"], 0, L["dens"], ft)
    px = (u0 * L["c"] * np.sin(2 * np.pi * (ix + dx) / L["size_x"])).astype(ft)
    z = np.zeros_like(px)
    ions = dict(ix=ix, iy=iy, dx=dx, dy=dy, px=z, py=z, pz=z)
    lecs = dict(ix=ix, iy=iy, dx=dx.copy(), dy=dy.copy(), px=px, py=z.copy(), pz=z.copy())
    return [(1.0, LANGMUIR_ION_MASS, ions), (-1.0, 1.0, lecs)]


def energies(sim, masses, dens, c):
    """Field energies sum |E|^2/2, sum |B|^2/2 over the interior and, per
    species, m c^2 sum (gamma - 1) / dens, computed with numpy from a state."""
    g = {n: sim.grid(n)[1:-1, 1:-1].astype(np.float64) for n in ["ex", "ey", "ez", "bx", "by", "bz"]}
    we = 0.5 * sum(float((g[n] ** 2).sum()) for n in ["ex", "ey", "ez"])
    wb = 0.5 * sum(float((g[n] ** 2).sum()) for n in ["bx", "by", "bz"])
    kin, mom = [], np.zeros(3)
    for sid, m in enumerate(masses):
        st = sim.species(sid)
        p = [st[k].astype(np.float64) for k in ("px", "py", "pz")]
        u2 = (p[0] ** 2 + p[1] ** 2 + p[2] ** 2) / (c * c)
        kin.append(m * c * c * float((u2 / (np.sqrt(1 + u2) + 1)).sum()) / dens)
        mom += [m * float(a.sum()) / dens for a in p]
    return we, wb, kin, mom
