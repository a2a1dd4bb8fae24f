"""Checkpoint and restore through the C ABI (marked gpu).

save_output (src/save/mod.rs:25-81) writes species 0 and the spatial fields;
a restart also needs every species and the persistent spectral B_wrk
(src/flds/mod.rs:31-33).  A run that is saved, loaded into a fresh context and
continued must equal the uninterrupted run: FP64 cells bit for bit and the
rest to 1e-11 (the FP64 deposit's LDS atomics sum in a varying order), and
every getter right after the load returns the saved state bit for bit."""
import os

import numpy as np
import pytest

from spectral_rs_amd import capi
from tests.helpers import shock_species

pytestmark = pytest.mark.gpu

KW = dict(size_x=256, size_y=32, delta=8, dens=8, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0, left_edge=1,
          sort_interval=3)
GRIDS = ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho", "dens")
SPECTRA = GRIDS + ("bxw", "byw", "bzw")


def shock(prec):
    g = capi.GpuSim(prec=prec, **KW)
    ft = np.float64 if prec == 8 else np.float32
    for q, m, vth, sd in [(1.0, 16.0, 0.0025, 3), (-1.0, 1.0, 0.01, 4)]:
        g.add_species(q, m, **shock_species(256, 32, 8, 8, 10.0, 10.0, vth, seed=sd, ft=ft))
    g.prestep()
    return g


def state(g):
    out = {f"g_{n}": g.grid(n) for n in GRIDS}
    out.update({f"s_{n}": g.spectrum(n) for n in SPECTRA})
    for sid in range(g.nspecies):
        for k, v in g.species_raw(sid).items():
            out[f"raw{sid}_{k}"] = v
    d = g.diag()
    out["steps"] = np.array(d["steps"])
    out["gauss_post_max"] = np.array(d["gauss_post_max"])
    return out


@pytest.mark.parametrize("prec", [4, 8])
def test_load_restores_every_buffer_bit_for_bit(prec, tmp_path):
    a = shock(prec)
    a.step(7)  # leaves a push pending: save runs it first
    path = os.fspath(tmp_path / "run.ckpt")
    a.checkpoint_save(path)
    b = capi.GpuSim(prec=prec, **KW)
    b.checkpoint_load(path)
    assert b.nspecies == 2
    sa, sb = state(a), state(b)
    assert sa.keys() == sb.keys()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert a.tuning() == b.tuning()


def test_fp64_continued_run_equals_the_uninterrupted_run(tmp_path):
    a = shock(8)
    a.step(7)
    path = os.fspath(tmp_path / "run.ckpt")
    a.checkpoint_save(path)
    b = capi.GpuSim(prec=8, **KW)
    b.checkpoint_load(path)
    del path
    for s in (a, b):
        s.step(11)  # crosses sorts at steps 9, 12, 15, 18
    for sid in range(2):
        x, y = a.species(sid), b.species(sid)
        for k in ("ix", "iy"):
            assert np.array_equal(x[k], y[k]), (sid, k)
        for k in ("dx", "dy", "px", "py", "pz"):
            assert np.max(np.abs(x[k] - y[k])) <= 1e-11 * max(np.max(np.abs(x[k])), 1e-30), (sid, k)
    for n in GRIDS:
        x, y = a.grid(n), b.grid(n)
        assert np.max(np.abs(x - y)) <= 1e-11 * max(np.max(np.abs(x)), 1e-30), n
    for n in ("bxw", "byw", "bzw"):
        x, y = a.spectrum(n), b.spectrum(n)
        assert np.max(np.abs(x - y)) <= 1e-11 * max(np.max(np.abs(x)), 1e-30), n
    assert a.diag()["steps"] == b.diag()["steps"] == 18


def test_fp32_continued_step_matches_to_summation_order(tmp_path):
    # One step after the load: the move reads only saved state, so cells and
    # offsets are bit-equal; the deposit windows sum in fixed point (order-
    # free) but are flushed to the grids with float atomics in a varying
    # order, so J, rho and dens agree to FP32 rounding of the summed
    # contributions: each species' current is up to |q| c times its density
    # (|v| < c), ~80 per cell here, while ion and electron currents cancel to
    # ~0.04 net, so the bound is a few FP32 ulps of c max(dens), not of |J|.
    # (The fields are judged through J for the same reason.)
    a = shock(4)
    a.step(5)
    path = os.fspath(tmp_path / "run.ckpt")
    a.checkpoint_save(path)
    b = capi.GpuSim(prec=4, **KW)
    b.checkpoint_load(path)
    for s in (a, b):
        s.step(1)
    for sid in range(2):
        x, y = a.species_raw(sid), b.species_raw(sid)
        for k in ("cell", "tag", "dx", "dy"):
            assert np.array_equal(x[k], y[k]), (sid, k)
    eps = float(np.finfo(np.float32).eps)
    dmax = float(np.max(a.grid("dens")))
    for n in ("jx", "jy", "jz", "rho", "dens"):
        x, y = a.grid(n).astype(np.float64), b.grid(n).astype(np.float64)
        scale = dmax * (KW["c"] if n.startswith("j") else 1.0)
        assert np.max(np.abs(x - y)) <= 8 * eps * scale, (n, np.max(np.abs(x - y)) / (eps * scale))


def test_load_rejects_mismatched_or_damaged_files(tmp_path):
    a = shock(8)
    a.step(2)
    path = os.fspath(tmp_path / "run.ckpt")
    a.checkpoint_save(path)
    other = dict(KW, size_y=64)
    with pytest.raises(capi.SpError, match="parameters differ"):
        capi.GpuSim(prec=8, **other).checkpoint_load(path)
    with pytest.raises(capi.SpError, match="parameters differ"):
        capi.GpuSim(prec=4, **KW).checkpoint_load(path)
    with pytest.raises(capi.SpError, match="without species"):
        a.checkpoint_load(path)
    raw = open(path, "rb").read()
    cut = os.fspath(tmp_path / "cut.ckpt")
    open(cut, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(capi.SpError, match="truncated"):
        capi.GpuSim(prec=8, **KW).checkpoint_load(cut)
    # a tag out of range: the first tag of species 0 follows the header, the
    # species record (charge, mass, n) and n cells
    n0 = a.count(0)
    hdr = len(raw) - (sum(a.count(s) * (4 + 4 + 5 * 8) + 24 for s in range(2)) + 11 * 34 * 258 * 8
                      + 14 * 32 * 256 * 16)
    off = hdr + 24 + 4 * n0
    bad = bytearray(raw)
    bad[off:off + 4] = np.int32(n0 + 5).tobytes()
    badp = os.fspath(tmp_path / "bad.ckpt")
    open(badp, "wb").write(bytes(bad))
    with pytest.raises(capi.SpError, match="out of range"):
        capi.GpuSim(prec=8, **KW).checkpoint_load(badp)
    with pytest.raises(capi.SpError, match="not a checkpoint"):
        junk = os.fspath(tmp_path / "junk.ckpt")
        open(junk, "wb").write(b"x" * 4096)
        capi.GpuSim(prec=8, **KW).checkpoint_load(junk)
    # the context that refused stays usable
    c = capi.GpuSim(prec=8, **KW)
    with pytest.raises(capi.SpError):
        c.checkpoint_load(badp)
    c.checkpoint_load(path)
    assert c.nspecies == 2
