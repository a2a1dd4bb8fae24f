"""FP64 diagnostics of the HIP library (sp_get_diag; no reference counterpart),
marked gpu.

  * the energy and momentum sums equal numpy sums over the same GPU state;
  * in a cold Langmuir oscillation, where field and kinetic energy trade
    places, their time-centred total stays within 1% (f64) / 2% (f32);
  * the pre-correction Gauss residual equals |k.E^n - ckc k.J + i rho/dens|
    computed with numpy FFTs from the grids the solve read (the E of the
    previous step, the filtered J and rho of this one); the post-correction
    residual is at round-off.
"""
import numpy as np
import pytest

from spectral_rs_amd import capi
from tests.helpers import LANGMUIR, TEST_SIM, energies, langmuir_species, shock_species

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [8, 4])
def test_langmuir_diag_sums_and_conservation(prec):
    ft = np.float64 if prec == 8 else np.float32
    g = capi.GpuSim(prec=prec, **LANGMUIR)
    sp = langmuir_species(ft=ft)
    for q, m, a in sp:
        g.add_species(q, m, **a)
    g.prestep()
    masses = [m for _, m, _ in sp]
    dens, c = LANGMUIR["dens"], LANGMUIR["c"]
    prev = g.diag()
    k0 = sum(prev["energy_kin"])
    tot = []
    for n in range(1, 129):
        g.step(1)
        d = g.diag()
        assert d["steps"] == n
        tot.append(d["energy_e"] + d["energy_b"] + 0.5 * (sum(d["energy_kin"]) + sum(prev["energy_kin"])))
        prev = d
        if n % 32 == 0:
            we, wb, kin, mom = energies(g, masses, dens, c)
            assert d["energy_e"] == pytest.approx(we, rel=1e-12, abs=1e-300)
            assert d["energy_b"] == pytest.approx(wb, rel=1e-12, abs=1e-30)
            for s in range(2):
                assert d["energy_kin"][s] == pytest.approx(kin[s], rel=1e-12, abs=1e-300)
            np.testing.assert_allclose(d["momentum"], mom, rtol=1e-10, atol=1e-12 * k0)
    tot = np.array(tot)
    assert (tot.max() - tot.min()) < (0.01 if prec == 8 else 0.02) * tot.mean()
    assert abs(tot.mean() - k0) < 0.03 * k0
    assert d["gauss_post_max"] < (1e-12 if prec == 8 else 1e-5)
    assert d["gauss_post"] <= d["gauss_post_max"]
    g.close()


def test_gauss_residual_matches_numpy_fft():
    g = capi.GpuSim(prec=8, **TEST_SIM)
    for q, sd in [(1.0, 1), (-1.0, 2)]:
        g.add_species(q, 1.0, **shock_species(24, 12, 5, 2, 3.0, 5.0, 1e-3, seed=sd, ft=np.float64))
    g.prestep()
    g.step(3)
    e_prev = {n: g.grid(n)[1:-1, 1:-1] for n in ("ex", "ey")}
    g.step(1)
    d = g.diag()
    nx, ny, dens = TEST_SIM["size_x"], TEST_SIM["size_y"], TEST_SIM["dens"]
    ckc = TEST_SIM["dt"] / dens  # src/flds/mod.rs:127
    f = {n: np.fft.fft2(g.grid(n)[1:-1, 1:-1]) for n in ("jx", "jy", "rho")}
    ex, ey = np.fft.fft2(e_prev["ex"]), np.fft.fft2(e_prev["ey"])
    kx = (2 * np.pi * np.fft.fftfreq(nx))[None, :]
    ky = (2 * np.pi * np.fft.fftfreq(ny))[:, None]
    # the curl terms of the E update are orthogonal to k and drop out
    res = kx * ex + ky * ey - ckc * (kx * f["jx"] + ky * f["jy"]) + 1j * f["rho"] / dens
    mask = np.ones((ny, nx), bool)
    mask[0, 0] = False
    mask[:, nx // 2] = False
    mask[ny // 2, :] = False
    rho2 = np.sum(np.abs(f["rho"][mask] / dens) ** 2)
    pre = np.sqrt(np.sum(np.abs(res[mask]) ** 2) / rho2)
    assert pre > 1e-3  # the reference's deposit does not conserve charge
    assert d["gauss_pre"] == pytest.approx(pre, rel=1e-9)
    assert d["gauss_post"] < 1e-13
    g.close()
