"""The C++ host program (mirror of the reference binary): config.toml parsing
and Config::new validation (src/lib.rs:69-117) run on the CPU; the end-to-end
run with the output writer (src/save/mod.rs:25-81) is a GPU test."""
import os
import subprocess

import numpy as np
import pytest

from spectral_rs_amd import build as B

REF_CONFIG = """[setup]
t_final = {t_final}

[params]
size_x = {size_x}
size_y = {size_y}
delta = {delta}
dt = 0.1 # In units 1/omega_pe
c = 3.0 # This is really c/omega_pe. Don't touch this.
dens = 2   # of prtls per species per cell
gamma_inj = {gamma_inj} # Speed of upstream flow
n_pass = 4 # Number of filter passes.

[output]
track_prtls = false # need to think of a better way of doing this.
track_interval = 100 # In timesteps
write_output = {write_output} # save output tsteps
output_interval = 1 # In timesteps
stride = 1 # Prtl downsampling
istep = {istep} # Flds downsampling
"""


def run_host(tmp_path, **kw):
    p = dict(t_final=500, size_x=2048, size_y=256, delta=64, gamma_inj=10.0, write_output="false", istep=1)
    p.update(kw)
    (tmp_path / "config.toml").write_text(REF_CONFIG.format(**p))
    if not os.path.exists(B.HOST):
        B.build_host()
    r = subprocess.run([B.HOST], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout


@pytest.mark.parametrize("kw,msg", [
    (dict(size_x=2047), "Number of cells in x direction must be even"),
    (dict(size_y=255), "Number of cells in y direction must be even"),
    (dict(delta=1024), "Delta must be less than 1/2 the size of size_x"),
    (dict(gamma_inj=1.0), "Lorentz factor must be greater than 1"),
    (dict(istep=3), "Number of cells in x & y direction must be divisible by 2"),
])
def test_config_errors(tmp_path, kw, msg):
    rc, out = run_host(tmp_path, **kw)
    assert rc == 1
    assert out.strip() == "Problem parsing arguments: " + msg


def test_missing_config(tmp_path):
    r = subprocess.run([B.HOST], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Could not open the config.toml file" in r.stdout


@pytest.mark.gpu
def test_host_run_writes_reference_output_layout(tmp_path):
    rc, out = run_host(tmp_path, t_final=3, size_x=48, size_y=24, delta=4, write_output="true")
    assert rc == 0, out
    for t in range(4):
        d = tmp_path / "output" / f"dat_{t:05d}"
        x = np.load(d / "x.npy")
        assert x.dtype == np.float32 and x.size == 2 * 40 * 24
        assert np.all((x >= 4) & (x < 46))
        gam = np.load(d / "gam.npy")
        assert np.all(gam >= 1.0)
        for f in ["Jx", "Jy", "Jz", "Ex", "Ey", "Ez", "Bx", "By", "Bz", "Dens"]:
            assert np.load(d / "flds" / f"{f}.npy").shape == (48 * 24,)


def write_cfg(d, **kw):
    p = dict(t_final=3, size_x=48, size_y=24, delta=4, gamma_inj=10.0, write_output="true", istep=1)
    p.update(kw)
    d.mkdir(parents=True, exist_ok=True)
    (d / "config.toml").write_text(REF_CONFIG.format(**p))
    return str(d / "config.toml")


def test_ensemble_members_must_share_t_final(tmp_path):
    a = write_cfg(tmp_path / "a", t_final=3)
    b = write_cfg(tmp_path / "b", t_final=4)
    if not os.path.exists(B.HOST):
        B.build_host()
    r = subprocess.run([B.HOST, a, b], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ensemble members must share t_final" in r.stdout


def test_sort_interval_must_not_be_negative(tmp_path):
    a = write_cfg(tmp_path)
    txt = open(a).read().replace("n_pass = 4", "n_pass = 4\nsort_interval = -1")
    open(a, "w").write(txt)
    if not os.path.exists(B.HOST):
        B.build_host()
    r = subprocess.run([B.HOST, a], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "`params.sort_interval` must be a non-negative integer" in r.stdout


@pytest.mark.gpu
def test_host_ensemble_writes_each_members_output(tmp_path):
    # two runs of different sizes (one autotuned) as one ensemble; each writes
    # the reference's layout beside its config file
    a = write_cfg(tmp_path / "a")
    b = write_cfg(tmp_path / "b", size_x=96, size_y=32, delta=8)
    txt = open(b).read().replace("n_pass = 4", "n_pass = 4\nautotune = true")
    open(b, "w").write(txt)
    if not os.path.exists(B.HOST):
        B.build_host()
    r = subprocess.run([B.HOST, a, b], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "autotuned" in r.stdout
    for d, n in [(tmp_path / "a", 2 * 40 * 24), (tmp_path / "b", 2 * 80 * 32)]:
        for t in range(4):
            x = np.load(d / "output" / f"dat_{t:05d}" / "x.npy")
            assert x.size == n
        assert np.load(d / "output" / "dat_00003" / "flds" / "Bz.npy").size in (48 * 24, 96 * 32)


def test_restart_is_refused_for_ensemble_members(tmp_path):
    a = write_cfg(tmp_path / "a")
    b = write_cfg(tmp_path / "b")
    txt = open(b).read().replace("[setup]", '[setup]\nrestart = "x.ckpt"')
    open(b, "w").write(txt)
    if not os.path.exists(B.HOST):
        B.build_host()
    r = subprocess.run([B.HOST, a, b], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "restart is for a single run" in r.stdout


@pytest.mark.gpu
def test_host_checkpoint_restart_continues_the_run(tmp_path):
    # FP64: a run that writes checkpoints every 5 steps, and a second run that
    # restarts from step 5; their outputs of steps 5..12 agree (cells exactly,
    # the rest to the FP64 deposit's summation order)
    extra = "n_pass = 4\nprecision = \"f64\"\nsort_interval = 3"
    a = write_cfg(tmp_path / "a", t_final=12)
    txt = open(a).read().replace("n_pass = 4", extra).replace("istep = 1", "istep = 1\ncheckpoint_interval = 5")
    open(a, "w").write(txt)
    if not os.path.exists(B.HOST):
        B.build_host()
    r = subprocess.run([B.HOST, a], cwd=tmp_path / "a", capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    ck = tmp_path / "a" / "output" / "checkpoint_00005.ckpt"
    assert ck.exists() and (tmp_path / "a" / "output" / "checkpoint_00010.ckpt").exists()
    b = write_cfg(tmp_path / "b", t_final=12)
    txt = open(b).read().replace("n_pass = 4", extra).replace("[setup]", f'[setup]\nrestart = "{ck}"')
    open(b, "w").write(txt)
    r = subprocess.run([B.HOST, b], cwd=tmp_path / "b", capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "restarted from" in r.stdout
    assert not (tmp_path / "b" / "output" / "dat_00004").exists()
    for t in (5, 9, 12):
        da, db = (tmp_path / d / "output" / f"dat_{t:05d}" for d in ("a", "b"))
        for f in ("x", "y"):
            xa, xb = np.load(da / f"{f}.npy"), np.load(db / f"{f}.npy")
            assert xa.dtype == np.float64
            assert np.array_equal(np.floor(xa), np.floor(xb)) and np.max(np.abs(xa - xb)) < 1e-11
        for f in ("u", "gam"):
            assert np.max(np.abs(np.load(da / f"{f}.npy") - np.load(db / f"{f}.npy"))) < 1e-11
        for f in ("Ex", "By", "Jx", "Dens"):
            fa, fb = np.load(da / "flds" / f"{f}.npy"), np.load(db / "flds" / f"{f}.npy")
            assert np.max(np.abs(fa - fb)) <= 1e-11 * max(np.max(np.abs(fa)), 1e-30), (t, f)
