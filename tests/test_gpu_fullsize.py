"""The bench workload at its full size (marked gpu): the 4096 x 1024 shock,
32 particles per cell per species, 266,338,304 particles, FP32, default
launch shapes, checked through properties that hold at any size:
  * particle counts are conserved and every particle stays inside the grid
    (the device asserts of src/prtls.rs:100-101, 155-170 stay silent);
  * every particle deposits a total TSC weight of 1, so the folded density
    grid sums to the particle count;
  * the Gauss correction leaves FP32 round-off (the pre-correction residual is
    the deposit's continuity error, far larger);
  * total energy (fields + kinetic) drifts by less than 1e-6 over 40 steps.
Measured on gfx950: density sum 3.4e-8 off, post-correction residual 4.4e-8
(pre / post 2.7e5), energy drift 8e-11; the bounds leave 10-30x."""
import sys

import numpy as np
import pytest

from spectral_rs_amd import capi

pytestmark = pytest.mark.gpu

C2 = dict(size_x=4096, size_y=1024, delta=16, dens=32, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)


def total_energy(d):
    return d["energy_e"] + d["energy_b"] + sum(d["energy_kin"])


def test_c2_shock_size_independent_properties():
    sim = capi.GpuSim(prec=4, sort_interval=20, left_edge=capi.SP_LEFT_REFLECT, **C2)
    for q, m, vth in [(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]:
        sim.init_species(q, m, vth, seed=0)
    per = C2["dens"] * (C2["size_x"] - 2 * C2["delta"]) * C2["size_y"]
    assert sim.count(0) == sim.count(1) == per == 133169152
    sim.prestep()
    e0 = total_energy(sim.diag())
    sim.step(40)
    sim.sync()  # raises SP_EINVARIANT if any particle left the grid or the slab
    assert sim.count(0) == sim.count(1) == per
    dens = sim.grid("dens")[1:-1, 1:-1].astype(np.float64).sum()
    assert abs(dens - 2 * per) <= 1e-6 * 2 * per, dens
    d = sim.diag()
    assert d["steps"] == 40
    assert d["gauss_post_max"] < 1e-6, d
    assert d["gauss_pre"] > 1e4 * d["gauss_post"], d
    e1 = total_energy(d)
    print(f"density sum rel err {dens / (2 * per) - 1:.3e}, gauss post max {d['gauss_post_max']:.3e}, "
          f"pre/post {d['gauss_pre'] / max(d['gauss_post'], 1e-300):.3e}, energy drift {e1 / e0 - 1:.3e}",
          file=sys.stderr)
    assert np.isfinite(e1) and abs(e1 - e0) <= 1e-6 * abs(e0), (e0, e1)
    sim.close()
