"""Pins the oracle's field functions to the reference's own unit-test vectors
(src/flds/field.rs:486-1543, src/flds/fft_2d.rs:478-952,
src/flds/wave_num.rs:49-943), to numpy, and to closed forms.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import E_TOL, TEST_SIM, ref_vectors

V = ref_vectors()
PRECS = [4, 8]


def sim(prec, **kw):
    p = dict(TEST_SIM)
    p.update(kw)
    return O.RefSim(prec=prec, **p)


@pytest.mark.parametrize("prec", PRECS)
def test_copy_to_spectral(prec):  # field.rs:486-569 (exact)
    s = sim(prec)
    s.set_grid("jx", V[("copy_to_spectral", "input")])
    s.copy_to_spectral("jx")
    sp = s.spectrum("jx")
    exp = V[("copy_to_spectral", "expected_output")].astype(s.ft)
    assert np.array_equal(sp.real.ravel(), exp)
    assert np.all(sp.imag == 0)


@pytest.mark.parametrize("prec", PRECS)
def test_copy_to_spatial(prec):  # field.rs:571-655 (exact)
    s = sim(prec)
    s.set_spectrum("jx", V[("copy_to_spatial", "input")].astype(s.ft))
    s.copy_to_spatial("jx")
    assert np.array_equal(s.grid("jx").ravel(), V[("copy_to_spatial", "expected_output")].astype(s.ft))


@pytest.mark.parametrize("prec", PRECS)
def test_ghost_deposit(prec):  # field.rs:656-723
    s = sim(prec)
    s.set_grid("jx", V[("ghost_deposit", "input")])
    s.deposit_ghosts("jx")
    assert np.max(np.abs(s.grid("jx").ravel() - V[("ghost_deposit", "expected_output")])) < E_TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_ghosts_update(prec):  # field.rs:725-795
    s = sim(prec)
    s.set_grid("jx", V[("ghosts_update", "input")])
    s.deposit_ghosts("jx")
    s.update_ghosts("jx")
    assert np.max(np.abs(s.grid("jx").ravel() - V[("ghosts_update", "expected_output")])) < E_TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_binomial_filter_reference_vector(prec):  # field.rs:797-1543
    s = sim(prec)
    s.set_grid("jx", V[("binomial_filter", "input")])
    for _ in range(TEST_SIM["n_pass"]):
        s.binomial_pass("jx")
    err = np.max(np.abs(s.grid("jx").ravel() - V[("binomial_filter", "expected_out.spatial")]))
    assert err < E_TOL[prec]


def _symbol_filter(a, npass):
    ny, nx = a.shape
    kx = 2 * np.pi * np.fft.fftfreq(nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny)
    sym = (np.cos(ky[:, None] / 2) ** 2 * np.cos(kx[None, :] / 2) ** 2) ** npass
    return np.real(np.fft.ifft2(np.fft.fft2(a) * sym))


@pytest.mark.parametrize("shape", [(12, 24), (16, 32), (10, 6)])
def test_binomial_closed_form_symbol(shape):
    # With periodic ghosts every pass multiplies mode k by cos^2(kx/2) cos^2(ky/2).
    ny, nx = shape
    rng = np.random.default_rng(3)
    a = rng.random((ny, nx))
    s = sim(8, size_x=nx, size_y=ny, delta=1)
    g = np.zeros((ny + 2, nx + 2))
    g[1:-1, 1:-1] = a
    s.set_grid("jx", g)
    s.update_ghosts("jx")
    for _ in range(4):
        s.binomial_pass("jx")
    out = s.grid("jx")[1:-1, 1:-1]
    assert np.max(np.abs(out - _symbol_filter(a, 4))) < 1e-13
    assert abs(out.sum() - a.sum()) < 1e-12  # the stencil conserves the sum


def test_binomial_stale_ghost_ring_first_pass():
    # The first pass reads the ring as it stands (deposit spill), not periodic
    # copies: a value parked in the top ghost row reaches row 1 of the output.
    s = sim(8, size_x=8, size_y=6, delta=1)
    g = np.zeros((8, 10))
    g[0, 3] = 16.0  # top ghost row, column 3
    s.set_grid("jx", g)
    s.binomial_pass("jx")
    out = s.grid("jx")
    assert out[1, 2] == 1.0 and out[1, 3] == 2.0 and out[1, 4] == 1.0
    assert out[1:-1, 1:-1].sum() == 4.0
    # the result's ring is periodic again (field.rs:422)
    assert np.array_equal(out[0, 1:-1], out[-2, 1:-1])
    assert np.array_equal(out[-1, 1:-1], out[1, 1:-1])


@pytest.mark.parametrize("prec", PRECS)
def test_wave_numbers(prec):  # wave_num.rs:49-943
    s = sim(prec)
    kx, ky, kn = s.wave_numbers()
    assert np.max(np.abs(kx.ravel() - V[("wave_nums_init", "k_x")])) < E_TOL[prec]
    assert np.max(np.abs(ky.ravel() - V[("wave_nums_init", "k_y")])) < E_TOL[prec]
    assert np.max(np.abs(kn.ravel() - V[("wave_nums_init", "k_norm")])) < E_TOL[prec]


def test_wave_numbers_closed_form():
    s = sim(8, size_x=16, size_y=10)
    kx, ky, kn = s.wave_numbers()
    fx = np.fft.fftfreq(16) * 2 * np.pi
    fy = np.fft.fftfreq(10) * 2 * np.pi
    fx[8] = np.pi  # Nyquist is positive (wave_num.rs:23)
    fy[5] = np.pi
    assert np.allclose(kx, np.broadcast_to(fx, (10, 16)), rtol=0, atol=1e-15)
    assert np.allclose(ky, np.broadcast_to(fy[:, None], (10, 16)), rtol=0, atol=1e-15)
    assert kn[0, 0] == 0.0


@pytest.mark.parametrize("prec", PRECS)
def test_forward_fft_1d(prec):  # fft_2d.rs:478-547 (two-sided)
    x = V[("forward_fft", "input")]
    x = x[:, 0] + 1j * x[:, 1]
    out = V[("forward_fft", "out")]
    out = out[:, 0] + 1j * out[:, 1]
    y = O.fft1d(x, prec=prec)
    assert np.max(np.abs(y.real - out.real)) < E_TOL[prec]
    assert np.max(np.abs(y.imag - out.imag)) < E_TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_fft_1d_roundtrip(prec):  # fft_2d.rs:549-600
    x = V[("fft_1d_roundtrip", "input")]
    x = x[:, 0] + 1j * x[:, 1]
    y = O.fft1d(O.fft1d(x, prec=prec), inverse=True, prec=prec)
    assert np.max(np.abs(y - x * x.size)) < E_TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_forward_fft2d(prec):  # fft_2d.rs:601-907, compared two-sided here
    x = V[("get_2d_input", "input")]
    s = sim(prec)
    s.set_spectrum("jx", (x[:, 0] + 1j * x[:, 1]).reshape(12, 24))
    s.fft2d("jx")
    out = V[("forward_fft2d", "out")]
    out = (out[:, 0] + 1j * out[:, 1]).reshape(12, 24)
    y = s.spectrum("jx")
    assert np.max(np.abs(y.real - out.real)) < E_TOL[prec]
    assert np.max(np.abs(y.imag - out.imag)) < E_TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
def test_fft_2d_roundtrip(prec):  # fft_2d.rs:929-952
    x = V[("get_2d_input", "input")]
    x = (x[:, 0] + 1j * x[:, 1]).reshape(12, 24)
    s = sim(prec)
    s.set_spectrum("jx", x)
    s.fft2d("jx")
    s.fft2d("jx", inverse=True)
    assert np.max(np.abs(s.spectrum("jx") - x)) < E_TOL[prec]


def test_regenerated_fft_inputs_match_reference():
    # The reference's FFT inputs are numpy default_rng(42) draws
    # (scratch/fft_gen.py:3-9); the extracted literals agree to 1e-15.
    a1 = np.random.default_rng(seed=42).random(24)
    a2 = np.random.default_rng(seed=42).random((12, 24))  # first row = the 1D input
    assert np.allclose(V[("forward_fft", "input")][:, 0], a1, atol=1e-15, rtol=0)
    assert np.allclose(V[("get_2d_input", "input")][:, 0].reshape(12, 24), a2, atol=1e-15, rtol=0)
    out = V[("forward_fft2d", "out")]
    assert np.allclose((out[:, 0] + 1j * out[:, 1]).reshape(12, 24), np.fft.fft2(a2), atol=2e-14, rtol=0)


@pytest.mark.parametrize("shape", [(8, 16), (64, 256), (12, 24), (6, 10), (32, 96), (7, 5)])
def test_fft2d_against_numpy(shape):
    ny, nx = shape
    rng = np.random.default_rng(7)
    x = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    s = sim(8, size_x=nx, size_y=ny, delta=1)
    s.set_spectrum("jx", x)
    s.fft2d("jx")
    ref = np.fft.fft2(x)
    assert np.max(np.abs(s.spectrum("jx") - ref)) / np.max(np.abs(ref)) < 1e-14
    s.fft2d("jx", inverse=True)
    assert np.max(np.abs(s.spectrum("jx") - x)) < 1e-13
    # f32 relative L2 error stays at the f32 rounding level
    s4 = sim(4, size_x=nx, size_y=ny, delta=1)
    s4.set_spectrum("jx", x)
    s4.fft2d("jx")
    e = np.linalg.norm(s4.spectrum("jx") - ref) / np.linalg.norm(ref)
    assert e < 5e-7 * np.log2(nx * ny)


def test_fft_closed_forms():
    s = sim(8, size_x=16, size_y=8, delta=1)
    d = np.zeros((8, 16), complex)
    d[0, 0] = 1.0
    s.set_spectrum("jx", d)
    s.fft2d("jx")
    assert np.max(np.abs(s.spectrum("jx") - 1.0)) < 1e-15  # delta -> constant
    y, x = np.mgrid[0:8, 0:16]
    m = np.exp(2j * np.pi * (3 * x / 16 + 2 * y / 8))
    s.set_spectrum("jx", m)
    s.fft2d("jx")
    sp = s.spectrum("jx")
    assert abs(sp[2, 3] - 128) < 1e-12  # single mode -> spike at (ky=2, kx=3)
    sp[2, 3] = 0
    assert np.max(np.abs(sp)) < 1e-12
