"""Launch shapes (sp_set_tuning), the autotuner (sp_autotune) and ensembles of
independent runs (sp_ensemble_*), marked gpu.

Every tuning computes the same physics: the knobs change only the launch
shapes, so FP64 runs with different tunings, and ensemble members against
stand-alone runs of the same inputs, agree to the order of the deposit
atomics (~1e-16, amplified to <1e-11 over a few steps), with bit-equal cells.
Ensembles are checked with the graph on (captures and replays, including an
odd step count that ends eagerly) and off, with uneven member sizes and
tunings, sorting at different phases."""
import numpy as np
import pytest

from spectral_rs_amd import capi
from tests.helpers import shock_species

pytestmark = pytest.mark.gpu


def rel(a, b):
    n = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / n) if n > 0 else float(np.max(np.abs(a)))


def check_same(a, b, tol=1e-11, nspecies=2):
    for name in ["ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho", "dens"]:
        assert rel(b.grid(name), a.grid(name)) < tol, name
    for sid in range(nspecies):
        sa, sb = a.species(sid), b.species(sid)
        assert np.array_equal(sa["ix"], sb["ix"]) and np.array_equal(sa["iy"], sb["iy"])
        for k in ["dx", "dy", "px", "py", "pz"]:
            assert rel(sb[k], sa[k]) < tol, k


def shock(nx, ny, dens, prec=8, sort_interval=3, delta=4, seed=0):
    kw = dict(size_x=nx, size_y=ny, delta=delta, dens=dens, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0,
              left_edge=capi.SP_LEFT_REFLECT)
    s = capi.GpuSim(prec=prec, sort_interval=sort_interval, **kw)
    ft = np.float64 if prec == 8 else np.float32
    for k, (q, m, vth) in enumerate([(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]):
        s.add_species(q, m, **shock_species(nx, ny, delta, dens, 10.0, 10.0, vth, seed=seed + k, ft=ft))
    return s


def test_tuning_defaults_and_validation():
    # FP32 default sort tiles: 8x4 at every density
    dense = shock(64, 16, 16, prec=4)
    assert (dense.tuning()["tile_w_log"], dense.tuning()["tile_h_log"]) == (3, 2)
    s = shock(64, 16, 2, prec=4)
    t = s.tuning()
    assert t == dict(chunk=2048, tile_w_log=3, tile_h_log=2, fft_rows=1, fft_cols=8, fft_radix_x=8, fft_radix_y=8)
    for bad, msg in [(dict(chunk=256), "chunk"), (dict(tile_w_log=0), "tile"), (dict(tile_h_log=6), "tile"),
                     (dict(fft_rows=0), "fft_rows"), (dict(fft_cols=65), "fft_cols"),
                     (dict(fft_radix_x=2), "radix")]:
        with pytest.raises(capi.SpError, match=msg):
            s.set_tuning(**bad)
    assert s.tuning() == t  # a rejected tuning changes nothing
    d = shock(64, 16, 2, prec=8)
    with pytest.raises(capi.SpError, match="1024 in FP64"):
        d.set_tuning(chunk=2048)


def test_fp32_chunk_sizes_give_the_same_cells_and_deposits():
    sims = [shock(128, 32, 8, prec=4, sort_interval=0) for _ in range(3)]
    for s, k in zip(sims, [512, 1024, 2048]):
        s.set_tuning(chunk=k)
        s.prestep()
    ref = sims[2]
    for s in sims[:2]:
        for sid in range(2):
            a, b = s.species(sid), ref.species(sid)
            for k in ["ix", "iy", "dx", "dy", "px"]:
                assert np.array_equal(a[k], b[k]), k
        # The FP32 deposit windows accumulate in fixed point with a step 2^-e
        # chosen per chunk from its largest per-cell count m (2^e <= 2^29 /
        # (25 m |q| c) for J, 2^29 / (25 m) for dens), so other chunks round
        # each contribution differently: per cell and species at most 25 m
        # steps, here with m <= 4 dens, for each of the two runs.
        m = 4 * 8
        lsb_j = 2.0 ** -(29 - np.floor(np.log2(25 * m * 10.0)))
        lsb_d = 2.0 ** -(29 - np.floor(np.log2(25 * m)))
        for name, lsb in [("jx", lsb_j), ("jy", lsb_j), ("jz", lsb_j), ("rho", lsb_d), ("dens", lsb_d)]:
            d = np.max(np.abs(s.grid(name).astype(np.float64) - ref.grid(name)))
            assert d <= 2 * 2 * 25 * m * lsb + 1e-6 * 80, (name, d)


def test_tunings_give_the_same_physics_fp64():
    a = shock(96, 48, 4)  # 96 = 2^5 3: mixed-radix plans
    b = shock(96, 48, 4)
    b.set_tuning(tile_w_log=2, tile_h_log=1, fft_rows=4, fft_cols=2, fft_radix_x=4, fft_radix_y=4)
    for s in (a, b):
        s.prestep()
        s.step(8)
    check_same(a, b)


def test_autotune_picks_a_valid_tuning_and_keeps_the_state():
    a = shock(128, 64, 4)
    b = shock(128, 64, 4)
    for s in (a, b):
        s.prestep()
        s.step(2)
    t = b.autotune(trials=2, use_cache=False)
    assert t == b.tuning()
    assert t["chunk"] == 1024 and 1 <= t["tile_w_log"] <= 5 and t["fft_radix_x"] in (4, 8)
    assert t["fft_rows"] * 128 <= 4096 and t["fft_cols"] * 64 <= 4096
    c = shock(128, 64, 4)
    assert c.autotune(trials=2, use_cache=True) == t  # kept for this configuration
    for s in (a, b):
        s.step(5)  # the autotuner ran on a scratch copy: b continues where it was
    check_same(a, b)
    f = shock(128, 64, 4, prec=4)
    tf = f.autotune(trials=2, use_cache=False)
    assert tf["chunk"] in (512, 1024, 2048)


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_members_match_stand_alone_runs(graph):
    sizes = [(64, 16, 2, 3), (128, 32, 4, 3), (96, 24, 2, 2), (64, 32, 8, 0)]
    alone = [shock(nx, ny, d, sort_interval=si, seed=3 * i) for i, (nx, ny, d, si) in enumerate(sizes)]
    members = [shock(nx, ny, d, sort_interval=si, seed=3 * i) for i, (nx, ny, d, si) in enumerate(sizes)]
    members[1].set_tuning(tile_w_log=2, tile_h_log=2, fft_rows=2, fft_radix_x=4)
    for s in alone + members:
        s.prestep()
    members[2].step(1)  # a member one step ahead: sorts at another phase
    alone[2].step(1)
    ens = capi.Ensemble(members, branches=3, unit=2)
    ens.set_graph(graph)
    for n in (7, 10):
        ens.step(n)
        ens.sync()
        for s in alone:
            s.step(n)
        for a, m in zip(alone, members):
            check_same(a, m)
    st = ens.stats()
    if graph:
        assert st["captures"] >= 1 and 1 <= st["kept"] <= 16
    else:
        assert st == dict(kept=0, captures=0)
    ens.close()
    for a, m in zip(alone, members):  # members work on their own again
        a.step(3)
        m.step(3)
        check_same(a, m)


def test_ensemble_rules():
    a, b = shock(64, 16, 2), shock(64, 16, 2)
    ens = capi.Ensemble([a, b], branches=2, unit=2)
    with pytest.raises(capi.SpError, match="already belongs"):
        capi.Ensemble([a], branches=1, unit=1)
    with pytest.raises(capi.SpError, match="ensemble"):
        a.set_graph(True)
    rc = capi.load().sp_destroy(a.h)
    assert rc == capi.SP_EINVAL
    with pytest.raises(capi.SpError, match="twice"):
        c = shock(64, 16, 2)
        capi.Ensemble([c, c])
    ens.close()
    a.close()
    b.close()
