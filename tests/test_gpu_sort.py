"""The tile sort, read in the device's own order (marked gpu).

The sort key of a particle is a pure function of its cell (tile of 2^twl x
2^thl cells, tile rows in serpentine order), so the keys are checked bit for
bit: after a sort with nothing moving, the keys read in device order are
non-decreasing, each tile holds exactly the particles whose cells lie in it,
and every particle keeps its values (found through its tag).  The order
inside a tile comes from atomics and is not checked.  The sort changes no
arithmetic: an FP64 run sorted every step equals an unsorted run (cells bit for
bit, fields to the deposit's summation order)."""
import numpy as np
import pytest

from oracle import oracle as O
from spectral_rs_amd import capi
from tests.helpers import shock_species

pytestmark = pytest.mark.gpu


def host_tile_key(ix, iy, nx, ny, twl, thl):
    """Serpentine tile index of each cell (the rule of the sort, restated)."""
    ntx = (nx + (1 << twl) - 1) >> twl
    nty = (ny + (1 << thl) - 1) >> thl
    tx = np.clip((ix - 1) >> twl, 0, ntx - 1)
    ty = np.clip((iy - 1) >> thl, 0, nty - 1)
    return ty * ntx + np.where(ty & 1, ntx - 1 - tx, tx), ntx * nty


def still_particles(n, nx, ny, delta, ft, seed):
    """Random cells and offsets at rest: the pre-loop move leaves every cell."""
    rng = np.random.default_rng(seed)
    z = np.zeros(n, ft)
    return dict(ix=rng.integers(delta + 2, nx - delta - 1, n), iy=rng.integers(1, ny + 1, n),
                dx=(rng.random(n) * 0.9 - 0.45).astype(ft), dy=(rng.random(n) * 0.9 - 0.45).astype(ft),
                px=z, py=z.copy(), pz=z.copy())


@pytest.mark.parametrize("prec,nx,ny,n,tile", [(4, 96, 48, 40000, (4, 1)), (4, 96, 48, 40000, (3, 2)),
                                               (4, 1536, 512, 400000, (3, 2)), (8, 96, 48, 40000, (3, 3)),
                                               (4, 144, 48, 20000, (5, 5))])
def test_device_order_is_sorted_by_bit_exact_tile_keys(prec, nx, ny, n, tile):
    # the 1536 x 512 grid has 24576 tiles of 8x4: the counter scan runs several
    # 64-counter steps per wave; 144 x 48 with 32x32 tiles has a partial tile
    # column and row
    kw = dict(size_x=nx, size_y=ny, delta=4, dens=4, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
    g = capi.GpuSim(prec=prec, sort_interval=1, **kw)
    g.set_tuning(chunk=1024 if prec == 8 else 2048, tile_w_log=tile[0], tile_h_log=tile[1])
    sp = still_particles(n, nx, ny, 4, g.ft, 5)
    g.add_species(-1.0, 1.0, **sp)
    g.prestep()  # sorts, then the (null) move and the deposit
    raw = g.species_raw(0)
    tag = raw["tag"]
    assert np.array_equal(np.sort(tag), np.arange(n))
    # every particle (through its tag) keeps its cell and values
    assert np.array_equal(raw["ix"], sp["ix"][tag]) and np.array_equal(raw["iy"], sp["iy"][tag])
    for k in ("dx", "dy", "px", "py", "pz"):
        assert np.array_equal(raw[k], sp[k][tag]), k
    key, ntiles = host_tile_key(raw["ix"], raw["iy"], nx, ny, *tile)
    assert np.all(np.diff(key) >= 0)
    want, _ = host_tile_key(sp["ix"], sp["iy"], nx, ny, *tile)
    assert np.array_equal(np.bincount(key, minlength=ntiles), np.bincount(want, minlength=ntiles))
    # before any sort the device order is the order of addition
    g2 = capi.GpuSim(prec=prec, sort_interval=0, **kw)
    g2.add_species(-1.0, 1.0, **sp)
    assert np.array_equal(g2.species_raw(0)["tag"], np.arange(n))


def test_sorted_and_unsorted_fp64_runs_agree():
    kw = dict(size_x=256, size_y=32, delta=8, dens=8, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0, left_edge=1)
    runs = {}
    for si in (0, 1):
        g = capi.GpuSim(prec=8, sort_interval=si, **kw)
        for q, m, vth, sd in [(1.0, 16.0, 0.0025, 3), (-1.0, 1.0, 0.01, 4)]:
            g.add_species(q, m, **shock_species(256, 32, 8, 8, 10.0, 10.0, vth, seed=sd, ft=np.float64))
        g.prestep()
        g.step(8)
        runs[si] = g
    a, b = runs[0], runs[1]
    for sid in range(2):
        sa, sb = a.species(sid), b.species(sid)
        for k in ("ix", "iy"):
            assert np.array_equal(sa[k], sb[k]), (sid, k)
        for k in ("dx", "dy", "px", "py", "pz"):
            assert np.max(np.abs(sa[k] - sb[k])) <= 1e-11 * max(np.max(np.abs(sa[k])), 1e-30), (sid, k)
    for name in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho", "dens"):
        x, y = a.grid(name), b.grid(name)
        assert np.max(np.abs(x - y)) <= 1e-11 * max(np.max(np.abs(x)), 1e-30), name
    # the unsorted run kept the order of addition, the sorted one did not
    assert np.array_equal(a.species_raw(1)["tag"], np.arange(a.count(1)))
    assert not np.array_equal(b.species_raw(1)["tag"], np.arange(b.count(1)))


def test_adaptive_sort_schedule_changes_only_the_order():
    """sp_set_sort_policy: a schedule that sorts at its longest interval (3
    steps; the rise of the slow-path fraction can never exceed 1) runs exactly
    the sorts it states, and neither it nor one driven by any rise of the
    slow-path fraction changes the arithmetic (FP64: cells bit for bit, the
    rest to the deposit's summation order)."""
    kw = dict(size_x=256, size_y=32, delta=8, dens=8, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0, left_edge=1)
    runs = {}
    for mode in ("fixed", "every3", "rise"):
        g = capi.GpuSim(prec=8, sort_interval=20, **kw)
        if mode == "every3":
            g.set_sort_policy(1, 3, 1.0)
        elif mode == "rise":
            g.set_sort_policy(1, 100, 0.0)
        for q, m, vth, sd in [(1.0, 16.0, 0.0025, 3), (-1.0, 1.0, 0.01, 4)]:
            g.add_species(q, m, **shock_species(256, 32, 8, 8, 10.0, 10.0, vth, seed=sd, ft=np.float64))
        g.prestep()
        g.step(10)
        runs[mode] = g
    # fixed: the pre-step and step 0; every3: the pre-step and steps 3, 6, 9
    assert runs["fixed"].diag()["sorts"] == 2
    assert runs["every3"].diag()["sorts"] == 4
    assert runs["rise"].diag()["sorts"] >= 1
    a = runs["fixed"]
    for b in (runs["every3"], runs["rise"]):
        for sid in range(2):
            sa, sb = a.species(sid), b.species(sid)
            for k in ("ix", "iy"):
                assert np.array_equal(sa[k], sb[k]), (sid, k)
            for k in ("dx", "dy", "px", "py", "pz"):
                assert np.max(np.abs(sa[k] - sb[k])) <= 1e-11 * max(np.max(np.abs(sa[k])), 1e-30), (sid, k)
        for name in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho", "dens"):
            x, y = a.grid(name), b.grid(name)
            assert np.max(np.abs(x - y)) <= 1e-11 * max(np.max(np.abs(x)), 1e-30), name
    # the arguments are validated; a context that never sorts has no schedule to adapt
    g = runs["fixed"]
    for bad in [(0, 3, 0.1), (4, 3, 0.1), (1, 3, -0.1), (1, 3, 1.5)]:
        with pytest.raises(capi.SpError):
            g.set_sort_policy(*bad)
    g.set_sort_policy(0, 0, 0.0)  # back to the fixed schedule
    u = capi.GpuSim(prec=8, sort_interval=0, **kw)
    with pytest.raises(capi.SpError):
        u.set_sort_policy(1, 3, 0.1)
