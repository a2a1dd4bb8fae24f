"""Shared test inputs: the reference's test fixture and seeded shock set-ups."""
import json
import os

import numpy as np

from spectral_rs_amd.inputs import momenta, positions, shock_species  # noqa: F401  (re-exported)

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
