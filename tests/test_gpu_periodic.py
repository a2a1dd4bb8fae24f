"""The cargo features as run-time flags, against the oracle (marked gpu).

* `periodic` (src/prtls.rs:136-143, Cargo.toml:29-31): x wraps like y instead
  of the wall; particles crossing x = 0 and x = Nx move bit for bit like the
  oracle's, and a multi-step FP64 run follows the oracle's FP64 run.
* the unchecked build (Cargo.toml:29-31, README.md:20): checked = 0 drops the
  asserts (src/prtls.rs:155-170) but not the arithmetic: the same cells and
  momenta as the oracle with its asserts off, and no SP_EINVARIANT where the
  checked build reports one.  Stencils that would leave the grid are still
  refused (the reference's slice indexing panics there in every build).
"""
import numpy as np
import pytest

from oracle import oracle as O
from spectral_rs_amd import capi
from tests.helpers import shock_species

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def edge_particles(n, nx, ny, ft, seed):
    """Random cells over the whole x range, a third of them in the first or
    last column, with momenta that carry many across x = 0 and x = Nx."""
    rng = np.random.default_rng(seed)
    ix = rng.integers(1, nx + 1, n)
    k = n // 3
    ix[:k] = np.where(rng.random(k) < 0.5, 1, nx)
    iy = rng.integers(1, ny + 1, n)
    d = lambda: (rng.random(n) - 0.5).astype(ft)
    p = lambda: (rng.standard_normal(n) * 6.0).astype(ft)
    return dict(ix=ix, iy=iy, dx=d(), dy=d(), px=p(), py=p(), pz=p())


@pytest.mark.parametrize("prec", [4, 8])
def test_periodic_x_mover_bit_exact(prec):
    nx, ny = 64, 32
    kw = dict(size_x=nx, size_y=ny, delta=4, dens=2, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0, periodic_x=1)
    g, r = capi.GpuSim(prec=prec, **kw), O.RefSim(prec=prec, **kw)
    sp = edge_particles(30000, nx, ny, g.ft, 9)
    g.add_species(-1.0, 1.0, **sp)
    r.add_species(-1.0, 1.0, **sp)
    g.move_and_deposit(0)
    r.move_and_deposit(0)
    a, b = g.species(0), r.species(0)
    for k in ["ix", "iy", "dx", "dy", "px", "py", "pz"]:
        assert np.array_equal(a[k], b[k]), k
    # crossings in both directions happened, and nothing reflected off a wall
    assert np.any((sp["ix"] == 1) & (a["ix"] == nx)) and np.any((sp["ix"] == nx) & (a["ix"] == 1))
    assert np.array_equal(np.sign(a["px"]), np.sign(sp["px"]))
    g.fold_all()
    r.fold_all()
    tol = 2e-5 if prec == 4 else 1e-12
    for name in ["jx", "jy", "jz", "rho", "dens"]:
        assert rel(g.grid(name), r.grid(name)) < tol, name
    assert abs(g.grid("dens")[1:-1, 1:-1].astype(np.float64).sum() - 30000) < (0.05 if prec == 4 else 1e-8)


def test_periodic_x_fp64_run_follows_the_oracle():
    # a drifting plasma in a box periodic in x and y: 20 steps of the fused
    # loop against the oracle's FP64 loop
    nx, ny, dens = 128, 32, 4
    kw = dict(size_x=nx, size_y=ny, delta=0, dens=dens, n_pass=4, dt=0.045, c=10.0, gamma_inj=3.0, periodic_x=1)
    g, r = capi.GpuSim(prec=8, sort_interval=5, **kw), O.RefSim(prec=8, **kw)
    for q, m, vth, sd in [(1.0, 16.0, 0.01, 1), (-1.0, 1.0, 0.05, 2)]:
        sp = shock_species(nx, ny, 0, dens, 10.0, 3.0, vth, seed=sd, ft=np.float64)
        g.add_species(q, m, **sp)
        r.add_species(q, m, **sp)
    for s in (g, r):
        s.prestep()
        s.step(20)
    for sid in range(2):
        a, b = g.species(sid), r.species(sid)
        assert np.array_equal(a["ix"], b["ix"]) and np.array_equal(a["iy"], b["iy"])
        assert rel(a["px"], b["px"]) < 1e-11
    for name in ["ex", "ey", "ez", "bx", "by", "bz", "jx", "rho"]:
        assert rel(g.grid(name), r.grid(name)) < 1e-10, name


@pytest.mark.parametrize("prec", [4, 8])
def test_unchecked_build_moves_like_the_oracle_without_asserts(prec):
    # one particle that crosses into cell delta (the checked build's
    # src/prtls.rs:163-170 assert) and one beyond the wall + 1
    # (src/prtls.rs:155-162), next to a shock slab
    kw = dict(size_x=64, size_y=16, delta=4, dens=2, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0)
    ft = np.float64 if prec == 8 else np.float32
    sp = shock_species(64, 16, 4, 2, 3.0, 5.0, 1e-3, seed=4, ft=ft)
    sp.pop("psa")
    wall = 64 + 1 - 4
    sp["ix"] = np.concatenate([sp["ix"], [5, wall + 2]])
    sp["iy"] = np.concatenate([sp["iy"], [8, 8]])
    for k, v in (("dx", [-0.45, 0.0]), ("dy", [0.0, 0.0]), ("px", [-30.0, 0.0]), ("py", [0.0, 0.0]),
                 ("pz", [0.0, 0.0])):
        sp[k] = np.concatenate([sp[k], np.array(v, ft)])
    for checked in (1, 0):
        g, r = capi.GpuSim(prec=prec, checked=checked, **kw), O.RefSim(prec=prec, checked=checked, **kw)
        g.add_species(-1.0, 1.0, **sp)
        r.add_species(-1.0, 1.0, **sp)
        if checked:
            with pytest.raises(O.OracleError, match="prtls.rs"):
                r.move_and_deposit(0)
            g.move_and_deposit(0)
            with pytest.raises(capi.SpError, match="prtls.rs"):
                g.sync()
            continue
        r.move_and_deposit(0)
        g.move_and_deposit(0)
        g.sync()  # nothing reported
        a, b = g.species(0), r.species(0)
        for k in ["ix", "iy", "dx", "dy", "px", "py", "pz"]:
            assert np.array_equal(a[k], b[k]), k
        assert a["ix"][-2] == 4  # in cell delta, where the checked build stops
        # a few full steps without asserts, pushes bit-exact on identical fields
        for s in (g, r):
            s.fold_all()
        for name in ["ex", "ey", "ez", "bx", "by", "bz"]:
            f = np.random.default_rng(len(name)).standard_normal((18, 66)).astype(ft) * 0.01
            g.set_grid(name, f)
            r.set_grid(name, f)
        g.boris_push(0)
        r.boris_push(0)
        a, b = g.species(0), r.species(0)
        for k in ["px", "py", "pz", "psa"]:
            assert np.array_equal(a[k], b[k]), k
