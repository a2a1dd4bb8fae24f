"""Step-loop graph mode (sp_set_graph), marked gpu: replayed periods give the
eager steps' results, checked after every replay with the program's own eager
work (diagnostics, flushes, a stand-alone filter that forces a re-capture)
between the replays.  FP64 on the reference fixture (src/lib.rs:219-244), so
the only difference left is the order of the deposit atomics (~1e-16)."""
import numpy as np
import pytest

from spectral_rs_amd import capi
from tests.helpers import TEST_SIM, shock_species

pytestmark = pytest.mark.gpu


def rel(a, b):
    n = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / n) if n > 0 else float(np.max(np.abs(a)))


def check_same(a, b):
    for name in ["ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "rho", "dens"]:
        assert rel(b.grid(name), a.grid(name)) < 1e-11, name
    for sid in range(2):
        sa, sb = a.species(sid), b.species(sid)
        assert np.array_equal(sa["ix"], sb["ix"]) and np.array_equal(sa["iy"], sb["iy"])
        for k in ["dx", "dy", "px", "py", "pz"]:
            assert rel(sb[k], sa[k]) < 1e-11, k


@pytest.mark.parametrize("sort_interval", [2, 0])
def test_graph_replays_match_eager_steps(sort_interval):
    sims = [capi.GpuSim(prec=8, sort_interval=sort_interval, **TEST_SIM) for _ in range(2)]
    eager, graph = sims
    for q, sd in [(1.0, 1), (-1.0, 2)]:
        sp = shock_species(24, 12, 5, 2, 3.0, 5.0, 1e-3, seed=sd, ft=np.float64)
        for s in sims:
            s.add_species(q, 1.0, **sp)
    period = 2 * sort_interval if sort_interval else 2
    for s in sims:
        s.prestep()
        s.step(period)  # the first step pushes nothing; the next period starts with a push pending
    graph.set_graph(True)
    for r in range(3):
        for s in sims:
            s.step(2 * period)  # capture (first round) and replays
            s.sync()
        check_same(eager, graph)
        da, db = eager.diag(), graph.diag()  # flushes the pending push: eager work between replays
        assert db["steps"] == da["steps"]
        assert db["energy_e"] == pytest.approx(da["energy_e"], rel=1e-11)
        assert db["gauss_post_max"] < 1e-12
        for s in sims:
            s.step(period)  # eager until the next period boundary with a push pending
        if r == 1:
            for s in sims:
                s.filter("ez", 1)  # swaps buffers: the next graph period is captured again
    check_same(eager, graph)
    for s in sims:
        s.close()
