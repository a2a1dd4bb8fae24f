"""GPU parity: every entry point of the HIP library against the CPU oracle on
identical inputs (marked gpu; run on an MI355X).

Bars (stated per test):
  * bit-exact: Boris push momenta and gamma, mover cells/offsets/momenta, ghost
    fold, binomial filter, spectral Maxwell update, GPU initialiser positions;
  * tolerance: deposits (float atomics sum in another order than the
    reference's serial loop), FFTs (own algorithm vs the oracle's), and
    everything downstream of them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from spectral_rs_amd import capi
from tests.helpers import TEST_SIM, ref_vectors, shock_species

pytestmark = pytest.mark.gpu

F32_EPS = np.finfo(np.float32).eps


def pair(prec=4, **kw):
    p = dict(TEST_SIM)
    p.update(kw)
    return capi.GpuSim(prec=prec, **p), O.RefSim(prec=prec, **p)


def random_fields(ny, nx, ft, seed, amp=(0.05, 0.05, 0.05, 1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    return {k: (rng.standard_normal((ny + 2, nx + 2)) * a).astype(ft)
            for k, a in zip(["ex", "ey", "ez", "bx", "by", "bz"], amp)}


def random_particles(n, nx, ny, delta, ft, seed, pscale=3.0):
    rng = np.random.default_rng(seed)
    ix = rng.integers(delta + 2, nx - delta - 1, n)
    iy = rng.integers(1, ny + 1, n)
    d = lambda: (rng.random(n) - 0.5).astype(ft)
    p = lambda: (rng.standard_normal(n) * pscale).astype(ft)
    return dict(ix=ix, iy=iy, dx=d(), dy=d(), px=p(), py=p(), pz=p())


def both(g, r, fn, *a):
    getattr(g, fn)(*a)
    getattr(r, fn)(*a)


@pytest.mark.parametrize("prec", [4, 8])
def test_boris_push_bit_exact(prec):
    g, r = pair(prec, size_x=96, size_y=48, delta=4)
    ft = g.ft
    for k, v in random_fields(48, 96, ft, 1).items():
        g.set_grid(k, v)
        r.set_grid(k, v)
    for q, m, sd in [(1.0, 16.0, 2), (-1.0, 1.0, 3)]:
        sp = random_particles(20000, 96, 48, 4, ft, sd)
        assert g.add_species(q, m, **sp) == r.add_species(q, m, **sp)
    for sid in range(2):
        both(g, r, "boris_push", sid)
        a, b = g.species(sid), r.species(sid)
        for k in ["px", "py", "pz", "psa"]:
            assert np.array_equal(a[k], b[k]), (sid, k, np.max(np.abs(a[k] - b[k])))


@pytest.mark.parametrize("prec", [4, 8])
def test_move_and_deposit_cells_bit_exact_and_deposit_within_tolerance(prec):
    nx, ny, delta = 64, 32, 4
    g, r = pair(prec, size_x=nx, size_y=ny, delta=delta)
    ft = g.ft
    sp = random_particles(30000, nx, ny, delta, ft, 7, pscale=6.0)
    g.add_species(-1.0, 1.0, **sp)
    r.add_species(-1.0, 1.0, **sp)
    both(g, r, "move_and_deposit", 0)
    a, b = g.species(0), r.species(0)
    for k in ["ix", "iy", "dx", "dy", "px", "py", "pz"]:
        assert np.array_equal(a[k], b[k]), k
    # per-cell bound: |J_gpu - J_ref| <= n_c * eps * sum|w q v|, with n_c the
    # number of contributions; sum|w q v| is the deposit of |q|, |p| at the
    # same mid-step positions (TSC weights are non-negative).
    absr = O.RefSim(prec=prec, **{**TEST_SIM, "size_x": nx, "size_y": ny, "delta": delta})
    mid = O.RefSim(prec=prec, **{**TEST_SIM, "size_x": nx, "size_y": ny, "delta": delta})
    mid.add_species(-1.0, 1.0, **sp)
    mid.update_position(0)
    mid.apply_bc(0)
    m = mid.species(0)
    absr.add_species(1.0, 1.0, m["ix"], m["iy"], m["dx"], m["dy"], np.abs(m["px"]), np.abs(m["py"]),
                     np.abs(m["pz"]), m["psa"])
    absr.deposit_current(0)
    eps = np.finfo(ft).eps
    cnt = 9 * 30000 / (nx * ny) * 8  # generous bound on contributions per cell
    # FP32 windows accumulate in fixed point with step 2^-e per chunk, where
    # 2^e <= 2^29 / (25 m |q| c) and m is the largest per-cell count of the
    # chunk (<= the global largest count used here); each of the <= 25 m
    # contributions to a cell is rounded to the step by floor(x + 1/2) in FP32,
    # off by at most one step.
    m = np.max(np.bincount((sp["iy"] * (nx + 2) + sp["ix"]).astype(np.int64)))
    if prec == 4:
        lsb_j = 2.0 ** -(29 - np.floor(np.log2(25 * m * 1.0 * TEST_SIM["c"])))
        lsb_d = 2.0 ** -(29 - np.floor(np.log2(25 * m)))
    else:
        lsb_j = lsb_d = 0.0
    for name in ["jx", "jy", "jz"]:
        A = absr.grid(name)
        d = np.abs(g.grid(name) - r.grid(name))
        bound = cnt * eps * A + 25 * m * 1.0 * lsb_j + eps * np.abs(r.grid(name))
        assert np.all(d <= bound + 1e-30), (name, float(np.max(d / (bound + 1e-30))))
    for name in ["rho", "dens"]:
        d = np.abs(g.grid(name) - r.grid(name))
        assert np.max(d) <= cnt * eps * np.max(r.grid("dens")) + 25 * m * 1.0 * lsb_d, name
    assert abs(g.grid("dens").astype(np.float64).sum() - 30000) < (0.05 if prec == 4 else 1e-8)


@pytest.mark.parametrize("prec", [4, 8])
def test_deposit_ghosts_bit_exact(prec):
    V = ref_vectors()
    g, r = pair(prec)
    for name in ["jx", "jy", "jz", "rho", "dens"]:
        g.set_grid(name, V[("ghost_deposit", "input")])
        r.set_grid(name, V[("ghost_deposit", "input")])
    g.fold_all()
    r.fold_all()
    for name in ["jx", "rho", "dens"]:
        assert np.array_equal(g.grid(name), r.grid(name))
    assert np.max(np.abs(g.grid("jx").ravel() - V[("ghost_deposit", "expected_output")])) < 1e-4


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("shape,passes", [((12, 24), 4), ((64, 256), 4), ((48, 96), 4), ((96, 384), 1),
                                          ((96, 384), 9), ((192, 576), 2), ((192, 576), 3), ((144, 288), 6)])
def test_binomial_filter_bit_exact_with_stale_ring(prec, shape, passes):
    # tiles of 32 x 64 cells (FP64: 16 x 64): the larger grids have tiles
    # whose first pass wraps around an edge and tiles whose frame (tile + the
    # passes' halo + the 1-cell ring) lies inside the padded grid, read into
    # LDS first, including frames that touch the ghost ring (one pass); grids
    # that are not multiples of the tile have partial tiles
    ny, nx = shape
    g, r = pair(prec, size_x=nx, size_y=ny, delta=2)
    a = np.random.default_rng(11).standard_normal((ny + 2, nx + 2)).astype(g.ft)  # non-periodic ring
    g.set_grid("jy", a)
    r.set_grid("jy", a)
    g.filter("jy", passes)
    for _ in range(passes):
        r.binomial_pass("jy")
    assert np.array_equal(g.grid("jy"), r.grid("jy"))


def test_binomial_filter_reference_vector():
    V = ref_vectors()
    g, _ = pair(4)
    g.set_grid("jx", V[("binomial_filter", "input")])
    g.filter("jx", 4)
    assert np.max(np.abs(g.grid("jx").ravel() - V[("binomial_filter", "expected_out.spatial")])) < 1e-4


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("shape", [(12, 24), (1024, 4096), (64, 256), (6, 48), (512, 2048), (768, 3072)])
def test_fft2d_against_fp64(prec, shape):
    ny, nx = shape
    g, _ = pair(prec, size_x=nx, size_y=ny, delta=2)
    rng = np.random.default_rng(ny + nx)
    x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(g.ct)
    g.set_spectrum("jx", x)
    g.fft2d("jx")
    ref = np.fft.fft2(x.astype(np.complex128))
    err = np.linalg.norm(g.spectrum("jx") - ref) / np.linalg.norm(ref)
    tol = (5e-7 if prec == 4 else 1e-15) * np.log2(nx * ny)
    assert err < tol, err
    g.fft2d("jx", inverse=True)
    back = g.spectrum("jx")
    err = np.linalg.norm(back - x) / np.linalg.norm(x)
    assert err < tol, err


@pytest.mark.parametrize("prec", [4, 8])
def test_fft2d_uneven_blocks(prec):
    # 3 rows / 3 columns per workgroup: the last block of each direction is
    # short and the index divisions are not shifts, on power-of-two lengths
    ny, nx = 64, 256
    g, _ = pair(prec, size_x=nx, size_y=ny, delta=2)
    g.set_tuning(fft_rows=3, fft_cols=3)
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(g.ct)
    g.set_spectrum("jx", x)
    g.fft2d("jx")
    ref = np.fft.fft2(x.astype(np.complex128))
    tol = (5e-7 if prec == 4 else 1e-15) * np.log2(nx * ny)
    assert np.linalg.norm(g.spectrum("jx") - ref) / np.linalg.norm(ref) < tol
    g.fft2d("jx", inverse=True)
    assert np.linalg.norm(g.spectrum("jx") - x) / np.linalg.norm(x) < tol


def test_fft2d_reference_vector():
    V = ref_vectors()
    g, _ = pair(4)
    x = V[("get_2d_input", "input")]
    g.set_spectrum("jx", (x[:, 0] + 1j * x[:, 1]).reshape(12, 24))
    g.fft2d("jx")
    out = V[("forward_fft2d", "out")]
    out = (out[:, 0] + 1j * out[:, 1]).reshape(12, 24)
    y = g.spectrum("jx")
    assert np.max(np.abs(y - out)) < 1e-4


@pytest.mark.parametrize("prec", [4, 8])
def test_spectral_update_bit_exact(prec):
    nx, ny = 64, 32
    g, r = pair(prec, size_x=nx, size_y=ny, delta=2, dens=3)
    rng = np.random.default_rng(5)
    for name in ["ex", "ey", "ez", "jx", "jy", "jz", "rho", "bxw", "byw", "bzw"]:
        s = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(g.ct)
        g.set_spectrum(name, s)
        r.set_spectrum(name, s)
    both(g, r, "spectral_update")
    for name in ["ex", "ey", "ez", "bx", "by", "bz", "bxw", "byw", "bzw", "jx", "jy"]:
        assert np.array_equal(g.spectrum(name), r.spectrum(name)), name


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


@pytest.mark.parametrize("prec", [4, 8])
def test_step_real_field_pipeline_spectra_and_update_match_oracle(prec):
    # With even sizes the step transforms its real fields two rows at a time
    # into half spectra, transforms columns kx = 0..Nx/2 (the getter fills the
    # other columns as their conjugate mirrors), projects only the
    # self-mirrored columns 0 and Nx/2, and updates the half plane (B_wrk on
    # both halves).  Checked here:
    #  * the J, rho spectra read back are exactly Hermitian and within FP rounding
    #    of numpy's FP64 fft2 of the interiors (rho / dens);
    #  * B_wrk equals, bit for bit, the oracle's update run on the spectra the
    #    step used: J, rho as stored, E rebuilt through the same pipeline in a
    #    second context (each E grid as that context's Jz, whose k = 0 mode the
    #    update keeps), and the initial B_wrk;
    #  * B_wrk stays exactly Hermitian.
    nx, ny, dens = 64, 32, 3
    g, r = pair(prec, size_x=nx, size_y=ny, delta=2, dens=dens, n_pass=0)
    ft = g.ft
    rng = np.random.default_rng(12)
    grids = {k: rng.standard_normal((ny + 2, nx + 2)).astype(ft) for k in ["ex", "ey", "ez", "jx", "jy", "jz", "rho"]}
    bw = {}
    for k in ["bxw", "byw", "bzw"]:
        x = rng.standard_normal((ny, nx)).astype(ft)
        bw[k] = np.fft.fft2(x).astype(g.ct)  # transform of a real field, nearly Hermitian
        bw[k] = 0.5 * (bw[k] + np.conj(bw[k][(-np.arange(ny)) % ny][:, (-np.arange(nx)) % nx]))
    for k, a in grids.items():
        g.set_grid(k, a)
    for k, a in bw.items():
        g.set_spectrum(k, a)
        r.set_spectrum(k, a)
    g.fields_update()
    mi, mj = (-np.arange(ny)) % ny, (-np.arange(nx)) % nx
    tol = 2e-6 if prec == 4 else 1e-14
    for k in ["jx", "jy", "jz", "rho"]:
        s = g.spectrum(k)
        assert np.array_equal(s[mi][:, mj], np.conj(s)), k
        x = grids[k][1:-1, 1:-1].astype(np.float64)
        ref = np.fft.fft2(x / dens if k == "rho" else x)
        if k in ("jx", "jy"):
            ref[0, 0] = 0.0  # mod.rs:106-108
        assert np.max(np.abs(s - ref)) <= tol * np.max(np.abs(ref)), k
        r.set_spectrum(k, s)
    aux = capi.GpuSim(prec=prec, **dict(TEST_SIM, size_x=nx, size_y=ny, delta=2, dens=dens, n_pass=0))
    for k in ["ex", "ey", "ez"]:
        for name in ["ex", "ey", "ez", "jx", "jy", "jz", "rho"]:
            aux.set_grid(name, np.zeros((ny + 2, nx + 2), ft))
        for name in ["bxw", "byw", "bzw"]:
            aux.set_spectrum(name, np.zeros((ny, nx), g.ct))
        aux.set_grid("jz", grids[k])
        aux.fields_update()
        r.set_spectrum(k, aux.spectrum("jz"))
    aux.close()
    r.spectral_update()
    for name in ["bxw", "byw", "bzw"]:
        s = g.spectrum(name)
        assert np.array_equal(s, r.spectrum(name)), name
        assert np.array_equal(s[mi][:, mj], np.conj(s)), name


@pytest.mark.parametrize("prec", [4, 8])
def test_half_plane_spectra_read_back_whole(prec):
    # The step's pipeline stores the J and rho spectra on the half plane
    # kx = 0..Nx/2 only; the getter fills the other columns as exact conjugate
    # mirrors, so what it returns is exactly Hermitian and the transform of the
    # (filtered) grid.  A spectrum set through the API is returned as set (not
    # mirrored), until the next update.
    nx, ny, dens = 64, 32, 3
    g, _ = pair(prec, size_x=nx, size_y=ny, delta=2, dens=dens, n_pass=2)
    rng = np.random.default_rng(21)
    for k in ["ex", "ey", "ez", "jx", "jy", "jz", "rho"]:
        g.set_grid(k, rng.standard_normal((ny + 2, nx + 2)).astype(g.ft))
    g.fields_update()
    mi, mj = (-np.arange(ny)) % ny, (-np.arange(nx)) % nx
    tol = 2e-6 if prec == 4 else 1e-14
    for k in ["jz", "jx", "jy", "rho"]:
        sp = g.spectrum(k)
        assert np.array_equal(sp[mi][:, mj], np.conj(sp)), k
        x = g.grid(k)[1:-1, 1:-1].astype(np.float64)
        ref = np.fft.fft2(x / dens if k == "rho" else x)
        if k in ("jx", "jy"):
            ref[0, 0] = 0.0  # mod.rs:106-108
        assert np.max(np.abs(sp - ref)) <= tol * np.max(np.abs(ref)), k
    x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(g.ct)
    g.set_spectrum("jz", x)
    assert np.array_equal(g.spectrum("jz"), x)
    g.fields_update()
    sp = g.spectrum("jz")
    assert np.array_equal(sp[mi][:, mj], np.conj(sp))
    g.close()


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("shape", [(128, 64), (54, 48), (96, 54)])
def test_fields_update_matches_oracle(prec, shape):
    # radix-2/4/8 plans; a radix-3 x plan; a radix-3 y plan
    nx, ny = shape
    g, r = pair(prec, size_x=nx, size_y=ny, delta=4, dens=4, n_pass=4)
    rng = np.random.default_rng(9)
    for name in ["ex", "ey", "ez", "jx", "jy", "jz", "rho"]:
        a = rng.standard_normal((ny + 2, nx + 2)).astype(g.ft)
        g.set_grid(name, a)
        r.set_grid(name, a)
    for _ in range(3):
        both(g, r, "fields_update")
    tol = 2e-5 if prec == 4 else 1e-12
    for name in ["ex", "ey", "ez", "bx", "by", "bz", "jx", "rho"]:
        assert rel(g.grid(name), r.grid(name)) < tol, (name, rel(g.grid(name), r.grid(name)))


def fp64_error_ratio(g, r32, r64, names):
    """Error of the GPU FP32 state and of the oracle's FP32 state, both against
    the FP64 (dprec) oracle run on the same inputs."""
    out = {}
    for name in names:
        ref = r64.grid(name)
        scale = np.max(np.abs(ref)) + 1e-30
        eg = np.max(np.abs(g.grid(name) - ref)) / scale
        er = np.max(np.abs(r32.grid(name) - ref)) / scale
        out[name] = (eg, er)
    return out


@pytest.mark.parametrize("sort_interval", [0, 2])
def test_full_step_on_reference_fixture_f64_bit_level(sort_interval):
    # build_test_sim (src/lib.rs:219-244), two species as in run() (lib.rs:124-126);
    # with sort_interval > 0 the particles are re-ordered by tile every 2 steps.
    g = capi.GpuSim(prec=8, sort_interval=sort_interval, **TEST_SIM)
    r = O.RefSim(prec=8, **TEST_SIM)
    for q, sd in [(1.0, 1), (-1.0, 2)]:
        sp = shock_species(24, 12, 5, 2, 3.0, 5.0, 1e-3, seed=sd, ft=np.float64)
        g.add_species(q, 1.0, **sp)
        r.add_species(q, 1.0, **sp)
    both(g, r, "prestep")
    both(g, r, "step", 5)
    for name in ["ex", "ey", "bz", "jx", "rho"]:
        assert rel(g.grid(name), r.grid(name)) < 1e-11, name
    for sid in range(2):
        a, b = g.species(sid), r.species(sid)
        assert np.array_equal(a["ix"], b["ix"])
        assert rel(a["px"], b["px"]) < 1e-11


def test_full_step_on_reference_fixture_f32_vs_dprec():
    # FP32 GPU vs the FP64 oracle, next to the oracle's own FP32 error.  In this
    # fixture ion and electron currents almost cancel, so |E| is ~1e-5 of the
    # individual contributions and FP32 summation order is visible.
    g, r = pair(4)
    r64 = O.RefSim(prec=8, **TEST_SIM)
    for q, sd in [(1.0, 1), (-1.0, 2)]:
        sp = shock_species(24, 12, 5, 2, 3.0, 5.0, 1e-3, seed=sd, ft=np.float32)
        g.add_species(q, 1.0, **sp)
        r.add_species(q, 1.0, **sp)
        r64.add_species(q, 1.0, **{k: v.astype(np.float64) if v.dtype == np.float32 else v for k, v in sp.items() if k != "psa"})
    for s in (g, r, r64):
        s.prestep()
    for sid in range(2):  # the pre-loop move depends only on the initial state
        a, b = g.species(sid), r.species(sid)
        for k in ["ix", "iy", "dx", "dy"]:
            assert np.array_equal(a[k], b[k]), k
    for s in (g, r, r64):
        s.step(5)
    for name, (eg, er) in fp64_error_ratio(g, r, r64, ["ex", "ey", "bz", "jx", "rho"]).items():
        assert eg <= 4 * er + 1e-5, (name, eg, er)


@pytest.mark.parametrize("sort_interval", [0, 7])
def test_shock_many_steps_fused_path_vs_dprec(sort_interval):
    # C1-shaped slab at reduced size: 256x32 cells, 8 ppc per species, mass ratio 16,
    # c = 10, dt = 0.045 (0.45 cells/step), 60 steps with the fused push+deposit pass.
    # The rear edge of the slab pulls a few electrons back past delta within
    # ~30 steps, which trips the reference's assert (src/prtls.rs:163-170);
    # both sides use the mirrored-wall left edge (left_edge=1) instead.
    kw = dict(size_x=256, size_y=32, delta=8, dens=8, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0, left_edge=1)
    g = capi.GpuSim(prec=4, sort_interval=sort_interval, **kw)
    r = O.RefSim(prec=4, **kw)
    r64 = O.RefSim(prec=8, **kw)
    for q, m, vth, sd in [(1.0, 16.0, 0.0025, 3), (-1.0, 1.0, 0.01, 4)]:
        sp = shock_species(256, 32, 8, 8, 10.0, 10.0, vth, seed=sd)
        for s in (g, r):
            s.add_species(q, m, **sp)
        r64.add_species(q, m, **{k: v.astype(np.float64) if v.dtype == np.float32 else v for k, v in sp.items() if k != "psa"})
    for s in (g, r, r64):
        s.prestep()
    for _ in range(3):
        for s in (g, r, r64):
            s.step(20)
        for name, (eg, er) in fp64_error_ratio(g, r, r64, ["ex", "bz", "rho"]).items():
            assert eg <= 4 * er + 1e-5, (name, eg, er)
    assert abs(g.grid("dens")[1:-1, 1:-1].astype(np.float64).sum() - 2 * 8 * 240 * 32) < 1.0


def test_gpu_initialiser_positions_and_momentum_moments():
    kw = dict(size_x=128, size_y=64, delta=8, dens=16, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0)
    g = capi.GpuSim(prec=4, **kw)
    sid = g.init_species(-1.0, 1.0, 0.01, 0)
    s = g.species(sid)
    ix, iy, dx, dy = O.init_positions(128, 64, 8, 16, prec=4)
    assert np.array_equal(s["ix"], ix) and np.array_equal(s["iy"], iy)
    assert np.array_equal(s["dx"], dx) and np.array_equal(s["dy"], dy)
    # Zenitani-boosted drifting Gaussian: <px> ~ gamma0 beta0 c, py, pz ~ N(0, (vth c)^2)
    c, g0 = 10.0, 10.0
    b0 = np.sqrt(1 - 1 / g0**2)
    assert np.mean(s["px"]) == pytest.approx(g0 * b0 * c, rel=2e-3)
    assert np.std(s["py"]) == pytest.approx(0.01 * c, rel=2e-2)
    assert abs(np.mean(s["pz"])) < 5e-3 * 0.01 * c * 10
    assert np.all(s["px"] > 0)


@pytest.mark.parametrize("nx,ny,n", [(96, 48, 40000), (1536, 512, 400000)])
def test_tile_sort_then_move_matches_the_oracle(nx, ny, n):
    # The sort changes the storage order only: after a sort and the pre-loop
    # move every particle (found through its tag) holds the oracle's values
    # bit for bit.  The device order and the tile keys themselves are checked
    # in test_gpu_sort.py.
    kw = dict(size_x=nx, size_y=ny, delta=4, dens=4, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0, left_edge=1)
    g = capi.GpuSim(prec=4, sort_interval=1, **kw)
    r = O.RefSim(prec=4, **kw)
    sp = random_particles(n, nx, ny, 4, np.float32, 21, pscale=4.0)
    g.add_species(-1.0, 1.0, **sp)
    r.add_species(-1.0, 1.0, **sp)
    both(g, r, "prestep")  # sorts, then moves
    a, b = g.species(0), r.species(0)
    for k in ["ix", "iy", "dx", "dy", "px", "py", "pz"]:
        assert np.array_equal(a[k], b[k]), k
