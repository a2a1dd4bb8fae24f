"""Precision-study variants of the FP32 path (sp_params.study), marked gpu:
each flag does what it states and nothing else."""
import numpy as np
import pytest

from spectral_rs_amd import capi
from tests.helpers import shock_species

pytestmark = pytest.mark.gpu

KW = dict(size_x=256, size_y=64, delta=8, dens=8, n_pass=4, dt=0.045, c=10.0, gamma_inj=10.0,
          left_edge=capi.SP_LEFT_REFLECT)


def shock(study):
    s = capi.GpuSim(prec=4, sort_interval=5, study=study, **KW)
    for q, m, vth, sd in [(1.0, 16.0, 0.0025, 1), (-1.0, 1.0, 0.01, 2)]:
        s.add_species(q, m, **shock_species(256, 64, 8, 8, 10.0, 10.0, vth, seed=sd))
    return s


def test_absolute_fp32_positions_lie_on_the_absolute_grid():
    s = shock(capi.SP_STUDY_ABS_POS32)
    s.prestep()
    s.step(3)
    for sid in range(2):
        st = s.species(sid)
        ix = st["ix"].astype(np.float32)
        for d, c in (("dx", ix), ("dy", st["iy"].astype(np.float32))):
            # an offset stored by an absolute-FP32 code: fl(fl(i + d) - i) == d
            back = (c + st[d]).astype(np.float32) - c
            assert np.array_equal(back.astype(np.float32), st[d]), d
    # the product path keeps finer offsets: most are off that grid
    m = shock(0)
    m.prestep()
    m.step(3)
    st = m.species(1)
    c = st["ix"].astype(np.float32)
    assert np.mean(((c + st["dx"]).astype(np.float32) - c) != st["dx"]) > 0.5


def test_fp32_twiddles_and_sums_change_only_rounding():
    base, tw = shock(0), shock(capi.SP_STUDY_TWIDDLE32 | capi.SP_STUDY_SUM32)
    for s in (base, tw):
        s.prestep()
        s.step(4)
    for name in ("ex", "bz", "jx"):
        a, b = base.grid(name), tw.grid(name)
        assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(a)), name
    da, db = base.diag(), tw.diag()
    for k in ("energy_e", "energy_b"):
        assert db[k] == pytest.approx(da[k], rel=1e-3), k
    assert db["energy_kin"][1] == pytest.approx(da["energy_kin"][1], rel=1e-4)
