"""Pins the oracle's particle and spectral functions, which no reference test
covers, to closed forms and invariants (CPU only).

* TSC weights (src/lib.rs:339-347): centre values, partition of unity,
  separable 1D form, and exact interpolation of linear fields (gather).
* update_position / apply_bc (src/prtls.rs:85-190): cell crossing, periodic y,
  the right wall and the left-edge assertion.
* deposit_current / calc_density (src/lib.rs:266-476): charge and current
  sums, stencil placement incl. ghost spill.
* boris_push (src/prtls.rs:262-443): uniform-E kick, an independent
  Rodrigues-rotation Boris step, |u| conservation in pure B.
* spectral update (src/flds/mod.rs:105-278): Gauss-law projection and the
  vacuum light-wave dispersion sin(w dt/2) = c |k| dt / 2.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import TEST_SIM, shock_species


def mk(prec=8, **kw):
    p = dict(TEST_SIM)
    p.update(kw)
    return O.RefSim(prec=prec, **p)


def one(s, ix, iy, dx, dy, px=0.0, py=0.0, pz=0.0, q=1.0, m=1.0):
    a = lambda v: np.atleast_1d(np.asarray(v, dtype=np.float64))
    return s.add_species(q, m, np.atleast_1d(ix), np.atleast_1d(iy), a(dx), a(dy), a(px), a(py), a(pz))


def tsc1d(d):
    return np.array([0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2])


def test_tsc_weights_centre_and_unity():
    s = mk(size_x=12, size_y=8, delta=2)
    one(s, 5, 4, 0.0, 0.0)
    s.calc_density(0)
    d = s.grid("dens")
    st = d[3:6, 4:7]
    assert np.allclose(st, np.outer([0.125, 0.75, 0.125], [0.125, 0.75, 0.125]), atol=0, rtol=0)
    assert d.sum() == 1.0


@pytest.mark.parametrize("dx,dy", [(0.3, -0.2), (-0.5, 0.49), (0.1, 0.4)])
def test_deposit_stencil_separable_and_placed(dx, dy):
    s = mk(size_x=12, size_y=8, delta=2)
    one(s, 5, 4, dx, dy, px=0.7, py=-0.4, pz=0.2, q=-1.0)
    s.deposit_current(0)
    s.calc_density(0)
    w = np.outer(tsc1d(dy), tsc1d(dx))  # rows = y, columns = x
    psa = np.sqrt(1 + (0.7**2 + 0.4**2 + 0.2**2) / 9.0)
    assert np.allclose(s.grid("dens")[3:6, 4:7], w, rtol=0, atol=1e-16)
    assert np.allclose(s.grid("dsty")[3:6, 4:7], -w, rtol=0, atol=1e-16)
    assert np.allclose(s.grid("jx")[3:6, 4:7], -w * 0.7 / psa, rtol=0, atol=1e-16)
    assert np.allclose(s.grid("jy")[3:6, 4:7], w * 0.4 / psa, rtol=0, atol=1e-16)
    assert np.allclose(s.grid("jz")[3:6, 4:7], -w * 0.2 / psa, rtol=0, atol=1e-16)
    assert abs(s.grid("dens").sum() - 1) < 1e-15


def test_deposit_spills_into_ghost_ring_then_folds():
    s = mk(size_x=12, size_y=8, delta=2)
    one(s, 3, 1, 0.0, 0.0)  # row 1 -> touches ghost row 0
    s.calc_density(0)
    g = s.grid("dens")
    assert g[0, 2:5].sum() == pytest.approx(0.125)
    s.deposit_ghosts("dens")
    g2 = s.grid("dens")
    assert g2[8, 2:5].sum() == pytest.approx(0.125)  # folded onto row Ny
    assert g2[1:-1, 1:-1].sum() == pytest.approx(1.0)


def test_shock_charge_and_current_sums():
    s = mk(prec=8)
    sp = shock_species(24, 12, 5, 2, 3.0, 5.0, 1e-3, seed=1, ft=np.float64)
    s.add_species(-1.0, 1.0, **sp)
    s.move_and_deposit(0)
    s.fold_all()
    n = sp["ix"].size
    inner = lambda g: s.grid(g)[1:-1, 1:-1].sum()
    assert inner("dens") == pytest.approx(n, rel=1e-14)
    assert inner("dsty") == pytest.approx(-n, rel=1e-14)
    st = s.species(0)  # psa unchanged by the mover
    assert inner("jx") == pytest.approx(-(st["px"] / st["psa"]).sum(), rel=1e-13)


def test_update_position_crossing():
    s = mk(size_x=12, size_y=8, delta=2)
    # c1 = (0.5 dt) / gamma ; dx += c1 px, one cell crossing each way
    one(s, 5, 4, 0.45, -0.45, px=2.0, py=-2.0)
    s.update_position(0)
    st = s.species(0)
    c1 = 0.05 / np.sqrt(1 + 8.0 / 9.0)
    assert st["ix"][0] == 6 and st["iy"][0] == 3
    assert st["dx"][0] == pytest.approx(0.45 + c1 * 2.0 - 1.0, abs=1e-15)
    assert st["dy"][0] == pytest.approx(-0.45 - c1 * 2.0 + 1.0, abs=1e-15)


def test_apply_bc_periodic_y_and_right_wall():
    s = mk(size_x=12, size_y=8, delta=2)  # wall = 12 + 1 - 2 = 11
    ix = np.array([5, 5, 11, 11, 12], np.int64)
    iy = np.array([0, 9, 4, 4, 4], np.int64)
    dx = np.array([0.1, 0.1, 0.2, -0.2, -0.3])
    s.add_species(1.0, 1.0, ix, iy, dx, np.zeros(5), np.array([1.0, 1, 1, 1, 1]), np.zeros(5), np.zeros(5))
    s.apply_bc(0)
    st = s.species(0)
    assert list(st["iy"]) == [8, 1, 4, 4, 4]
    # ix == wall, dx >= 0: reflected (px, dx sign flipped); dx < 0: untouched
    assert st["px"][2] == -1.0 and st["dx"][2] == -0.2
    assert st["px"][3] == 1.0 and st["dx"][3] == -0.2
    # ix > wall: back one cell, dx + 1, px flipped
    assert st["ix"][4] == 11 and st["dx"][4] == pytest.approx(0.7) and st["px"][4] == -1.0


def test_left_edge_assertion_and_reflect_rule():
    s = mk(size_x=12, size_y=8, delta=2)
    one(s, 2, 4, 0.0, 0.0)  # ix <= delta trips prtls.rs:163-170
    with pytest.raises(O.OracleError, match="prtls.rs:163"):
        s.apply_bc(0)
    s2 = mk(size_x=12, size_y=8, delta=2, left_edge=1)
    s2.add_species(1.0, 1.0, np.array([2, 3, 3]), np.array([4, 4, 4]), np.array([0.2, -0.1, 0.1]),
                   np.zeros(3), np.array([-1.0, -1.0, -1.0]), np.zeros(3), np.zeros(3))
    s2.apply_bc(0)
    st = s2.species(0)
    assert list(st["ix"]) == [3, 3, 3]
    assert st["dx"][0] == pytest.approx(-0.8) and st["px"][0] == 1.0
    assert st["dx"][1] == pytest.approx(0.1) and st["px"][1] == 1.0
    assert st["dx"][2] == pytest.approx(0.1) and st["px"][2] == -1.0


def _field_const(s, name, f):
    g = np.zeros((s.ny + 2, s.nx + 2))
    g[:] = f
    s.set_grid(name, g)


def test_gather_is_exact_for_linear_fields():
    s = mk(size_x=12, size_y=8, delta=2)
    yy, xx = np.mgrid[0:10, 0:14].astype(np.float64)
    lin = {"ex": 0.3 + 0.01 * xx, "ey": -0.2 + 0.02 * yy, "ez": 0.01 * xx - 0.03 * yy,
           "bx": 0.0 * xx, "by": 0.0 * xx, "bz": 0.0 * xx}
    for k, v in lin.items():
        s.set_grid(k, v)
    dx, dy = 0.31, -0.27
    one(s, 5, 4, dx, dy, q=1.0, m=1.0)
    s.boris_push(0)
    st = s.species(0)
    beta = 0.5 * 0.1  # (q/m) * 0.5 * dt
    x, y = 5 + dx, 4 + dy
    e = np.array([0.3 + 0.01 * x, -0.2 + 0.02 * y, 0.01 * x - 0.03 * y])
    # B = 0: p' = (p + beta E) + beta E
    assert np.allclose([st["px"][0], st["py"][0], st["pz"][0]], 2 * beta * e, rtol=1e-14, atol=1e-16)


def _rodrigues_boris(p, E, B, q, m, dt, c):
    beta = q / m * 0.5 * dt
    alpha = q / m * 0.5 * dt / c
    um = p + beta * E
    g = np.sqrt(1 + um @ um / c**2)
    t = alpha * B / g
    tn = np.linalg.norm(t)
    k = t / tn
    th = -2 * np.arctan(tn)  # Boris rotates u- about t by -2 atan|t|
    up = um * np.cos(th) + np.cross(k, um) * np.sin(th) + k * (k @ um) * (1 - np.cos(th))
    pn = up + beta * E
    return pn, np.sqrt(1 + pn @ pn / c**2)


@pytest.mark.parametrize("q", [1.0, -1.0])
def test_boris_matches_rodrigues_rotation(q):
    s = mk(size_x=12, size_y=8, delta=2)
    E = np.array([0.02, -0.05, 0.03])
    B = np.array([0.4, 0.9, -1.3])
    for k, v in zip(["ex", "ey", "ez", "bx", "by", "bz"], list(E) + list(B)):
        _field_const(s, k, v)
    p = np.array([1.5, -0.7, 0.4])
    one(s, 5, 4, 0.1, 0.2, *p, q=q, m=2.0)
    s.boris_push(0)
    st = s.species(0)
    pn, g = _rodrigues_boris(p, E, B, q, 2.0, 0.1, 3.0)
    assert np.allclose([st["px"][0], st["py"][0], st["pz"][0]], pn, rtol=1e-13, atol=1e-15)
    assert st["psa"][0] == pytest.approx(g, rel=1e-14)


def test_pure_b_preserves_momentum_magnitude():
    s = mk(size_x=12, size_y=8, delta=2)
    _field_const(s, "bz", 2.0)
    _field_const(s, "bx", 0.5)
    p = np.array([2.0, 1.0, -0.5])
    one(s, 5, 4, 0.0, 0.0, *p)
    for _ in range(50):
        s.boris_push(0)
    st = s.species(0)
    pn = np.array([st["px"][0], st["py"][0], st["pz"][0]])
    assert np.linalg.norm(pn) == pytest.approx(np.linalg.norm(p), rel=1e-13)
    assert not np.allclose(pn, p)


def test_gauss_projection():
    # After the correction k.E = -i rho on every non-zero mode that is not on a
    # Nyquist line (flds/mod.rs:185-201 then 230-266 zero those lines).
    nx, ny = 16, 8
    s = mk(size_x=nx, size_y=ny, delta=2, dens=1)
    rng = np.random.default_rng(5)
    for name in ["ex", "ey", "ez", "jx", "jy", "jz", "dsty"]:
        s.set_spectrum(name, np.fft.fft2(rng.standard_normal((ny, nx))))
    s.spectral_update()
    kx, ky, _ = s.wave_numbers()
    res = kx * s.spectrum("ex") + ky * s.spectrum("ey") + 1j * s.spectrum("dsty")
    mask = np.ones((ny, nx), bool)
    mask[0, 0] = False
    mask[:, nx // 2] = False
    mask[ny // 2, :] = False
    assert np.max(np.abs(res[mask])) < 1e-12
    assert s.spectrum("jx")[0, 0] == 0 and s.spectrum("jy")[0, 0] == 0


def test_vacuum_light_wave_dispersion():
    # Ez = cos(kx x) standing wave, B = 0 at t=0.  Leapfrog with E^1 = E^0 gives
    # E^n = cos((n - 1/2) w dt) / cos(w dt / 2) with sin(w dt/2) = c k dt / 2.
    nx, ny = 32, 4
    c, dt = 3.0, 0.1
    s = mk(size_x=nx, size_y=ny, delta=2, c=c, dt=dt, n_pass=0)
    m = 3
    k = 2 * np.pi * m / nx
    x = np.arange(nx)
    g = np.zeros((ny + 2, nx + 2))
    g[1:-1, 1:-1] = np.cos(k * x)[None, :]
    s.set_grid("ez", g)
    s.update_ghosts("ez")
    w = 2 / dt * np.arcsin(c * k * dt / 2)
    for n in range(1, 40):
        s.fields_update()
        amp = s.grid("ez")[2, 1]
        assert amp == pytest.approx(np.cos((n - 0.5) * w * dt) / np.cos(w * dt / 2), abs=1e-11)


def test_f32_and_f64_oracles_agree_on_a_shock_step():
    out = {}
    for prec, ft in [(4, np.float32), (8, np.float64)]:
        s = mk(prec=prec)
        for q, sd in [(1.0, 1), (-1.0, 2)]:
            s.add_species(q, 1.0, **shock_species(24, 12, 5, 2, 3.0, 5.0, 1e-3, seed=sd, ft=ft))
        s.prestep()
        s.step(3)
        out[prec] = (s.grid("ex"), s.species(1))
    e4, e8 = out[4][0], out[8][0]
    assert np.max(np.abs(e4 - e8)) <= 1e-4 * max(np.max(np.abs(e8)), 1e-3)
    assert np.array_equal(out[4][1]["ix"], out[8][1]["ix"])


def test_langmuir_energy_exchange_and_conservation():
    # Cold plasma oscillation at omega_pe = 1 (J/dens normalisation, m = |q| = 1):
    # field and kinetic energy trade places every quarter period, and the total
    # with the kinetic energy centred in time (p lives on half steps) stays
    # within 1% over two plasma periods.
    from tests.helpers import LANGMUIR, energies, langmuir_species
    s = O.RefSim(prec=8, **LANGMUIR)
    sp = langmuir_species()
    for q, m, a in sp:
        s.add_species(q, m, **a)
    s.prestep()
    masses = [m for _, m, _ in sp]
    prev = energies(s, masses, LANGMUIR["dens"], LANGMUIR["c"])
    k0 = sum(prev[2])
    tot, we = [], []
    for n in range(128):
        s.step(1)
        e = energies(s, masses, LANGMUIR["dens"], LANGMUIR["c"])
        tot.append(e[0] + e[1] + 0.5 * (sum(e[2]) + sum(prev[2])))
        we.append(e[0])
        prev = e
    tot, we = np.array(tot), np.array(we)
    assert we.max() > 0.9 * k0 and we.min() < 0.05 * k0
    assert (tot.max() - tot.min()) < 0.01 * tot.mean()
    assert abs(tot.mean() - k0) < 0.03 * k0
    # energy period = half the plasma period 2 pi / (omega_pe dt) = 62.8 steps
    up = np.flatnonzero((we[1:] > 0.5 * k0) & (we[:-1] <= 0.5 * k0))
    assert np.diff(up).mean() == pytest.approx(np.pi / LANGMUIR["dt"], rel=0.05)
