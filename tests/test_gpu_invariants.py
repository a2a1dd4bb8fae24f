"""The reference's asserts on the GPU (marked gpu).  A particle that crosses
the left edge (src/prtls.rs:163-170) or stands beyond the wall
(src/prtls.rs:155-162) sets a device flag in the particle pass, which exits
normally; the next synchronising call returns SP_EINVARIANT with the text of
the assert, where the oracle raises the same assert.  The flag is cleared by
that report.  Also an empty species next to a populated one."""
import numpy as np
import pytest

from oracle import oracle as O
from spectral_rs_amd import capi

pytestmark = pytest.mark.gpu

KW = dict(size_x=64, size_y=16, delta=4, dens=2, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0)


def one(ix, dx, px, ft):
    a = lambda v: np.array([v], ft)
    return dict(ix=np.array([ix]), iy=np.array([8]), dx=a(dx), dy=a(0.0), px=a(px), py=a(0.0), pz=a(0.0))


@pytest.mark.parametrize("prec", [4, 8])
def test_left_edge_assert(prec):
    g, r = capi.GpuSim(prec=prec, **KW), O.RefSim(prec=prec, **KW)
    ft = g.ft
    # |p| = 30 at c = 3: v ~ -c, so the first half move (0.5 dt v = -0.15)
    # takes dx from -0.45 across the cell edge into cell delta
    sp = one(KW["delta"] + 1, -0.45, -30.0, ft)
    g.add_species(-1.0, 1.0, **sp)
    r.add_species(-1.0, 1.0, **sp)
    with pytest.raises(O.OracleError, match="prtls.rs:163-170"):
        r.move_and_deposit(0)
    g.move_and_deposit(0)
    with pytest.raises(capi.SpError, match="prtls.rs:163-170"):
        g.sync()
    g.sync()  # reported once: the flag is clear again


@pytest.mark.parametrize("prec", [4, 8])
def test_wall_assert(prec):
    g, r = capi.GpuSim(prec=prec, **KW), O.RefSim(prec=prec, **KW)
    wall = KW["size_x"] + 1 - KW["delta"]  # src/prtls.rs:153
    sp = one(wall + 2, 0.0, 0.0, g.ft)    # at rest beyond wall + 1
    g.add_species(-1.0, 1.0, **sp)
    r.add_species(-1.0, 1.0, **sp)
    with pytest.raises(O.OracleError, match="prtls.rs:155-162"):
        r.move_and_deposit(0)
    g.move_and_deposit(0)
    with pytest.raises(capi.SpError, match="prtls.rs:155-162"):
        g.sync()


def test_left_edge_reflect_raises_nothing():
    # the mirrored-wall left edge (left_edge=1) turns the same particle back
    g = capi.GpuSim(prec=8, left_edge=capi.SP_LEFT_REFLECT, **KW)
    g.add_species(-1.0, 1.0, **one(KW["delta"] + 1, -0.45, -30.0, g.ft))
    g.move_and_deposit(0)
    g.sync()
    assert g.species(0)["ix"][0] > KW["delta"]


def test_empty_species_changes_nothing():
    rng = np.random.default_rng(3)
    n = 2000
    sp = dict(ix=rng.integers(6, 58, n), iy=rng.integers(1, 17, n), dx=rng.random(n) - 0.5, dy=rng.random(n) - 0.5,
              px=rng.standard_normal(n), py=rng.standard_normal(n), pz=rng.standard_normal(n))
    empty = {k: v[:0] for k, v in sp.items()}
    runs = []
    for with_empty in (False, True):
        g = capi.GpuSim(prec=8, sort_interval=2, left_edge=capi.SP_LEFT_REFLECT, **KW)
        if with_empty:
            g.add_species(1.0, 16.0, **empty)
        g.add_species(-1.0, 1.0, **sp)
        g.prestep()
        g.step(3)
        g.sync()
        runs.append((g, 1 if with_empty else 0))
    (a, ia), (b, ib) = runs
    assert b.count(0) == 0 and b.species(0)["ix"].size == 0
    sa, sb = a.species(ia), b.species(ib)
    assert np.array_equal(sa["ix"], sb["ix"]) and np.array_equal(sa["iy"], sb["iy"])
    for k in ["dx", "dy", "px", "py", "pz"]:
        assert np.max(np.abs(sa[k] - sb[k])) <= 1e-11 * max(1.0, np.max(np.abs(sb[k]))), k
    for name in ["ex", "bz", "jx", "rho"]:
        x, y = a.grid(name), b.grid(name)
        assert np.max(np.abs(x - y)) <= 1e-11 * max(1.0, np.max(np.abs(y))), name
