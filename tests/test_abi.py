"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/spectral_rs.h declares, and rejects the configurations that
Config::new rejects (src/lib.rs:75-113) before touching a GPU."""
import ctypes as ct
import os
import re

import pytest

from spectral_rs_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "spectral_rs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"^\s*(?:int|const char\*)\s+(sp_\w+)\s*\(", txt, flags=re.M)


def test_header_and_binding_agree():
    assert header_symbols() == capi.SYMBOLS


def test_library_exports_every_declared_symbol():
    L = capi.load()
    for name in header_symbols():
        assert hasattr(L, name), name


def _params(**kw):
    p = dict(size_x=24, size_y=12, delta=5, dens=2, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0)
    p.update(kw)
    return capi.SpParams(p["size_x"], p["size_y"], p["delta"], p["dens"], p["n_pass"], p["dt"], p["c"],
                         p["gamma_inj"], p.get("precision", 0), 0, 1, p.get("left_edge", 0), 0, p.get("study", 0))


@pytest.mark.parametrize("kw,msg", [
    (dict(size_x=25), "Number of cells in x direction must be even"),
    (dict(size_y=13), "Number of cells in y direction must be even"),
    (dict(size_x=25, size_y=13), "Number of cells in x & y direction must be even"),
    (dict(delta=12), "Delta must be less than 1/2 the size of size_x"),
    (dict(gamma_inj=1.0), "Lorentz factor must be greater than 1"),
    (dict(size_x=20), "2^a 3^b"),
    (dict(size_x=16384), "at most"),
    (dict(precision=7), "unknown precision"),
    (dict(left_edge=9), "unknown left_edge"),
    (dict(study=8), "unknown study flag"),
    (dict(study=1, precision=1), "study variants are of the FP32 path"),
])
def test_config_validation(kw, msg):
    L = capi.load()
    h = ct.c_void_p()
    p = _params(**kw)
    rc = L.sp_create(ct.byref(p), ct.byref(h))
    assert rc == capi.SP_EINVAL
    assert msg in L.sp_last_error(None).decode()


def test_null_arguments_are_rejected():
    L = capi.load()
    assert L.sp_create(None, None) == capi.SP_EINVAL
    assert L.sp_sync(None) == capi.SP_EINVAL
    assert L.sp_step(None, 1) == capi.SP_EINVAL
