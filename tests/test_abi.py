"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/spectral_rs.h declares, and rejects the configurations that
Config::new rejects (src/lib.rs:75-113) before touching a GPU."""
import ctypes as ct
import os
import re
import subprocess

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


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof of a header struct, from a C compiler."""
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("%zu\\n", offsetof({struct}, {f}));\n' for f in fields)
    src.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "spectral_rs.h"\n'
                   f'int main(void) {{\n  printf("%zu\\n", sizeof({struct}));\n{body}  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    return out[0], out[1:]


@pytest.mark.parametrize("struct,cls", [("sp_params", capi.SpParams), ("sp_diag", capi.SpDiag),
                                        ("sp_tuning", capi.SpTuning)])
def test_struct_layouts_match_the_header(struct, cls, tmp_path):
    names = [n for n, _ in cls._fields_]
    size, offs = _c_layout(struct, names, tmp_path)
    assert ct.sizeof(cls) == size
    assert [getattr(cls, n).offset for n in names] == offs


def test_documented_rust_binding_mirrors_sp_params():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = re.search(r"pub struct SpParams \{(.*?)\}", txt, flags=re.S).group(1)
    rust = re.findall(r"pub (\w+): (\[i32; \d+\]|i32|f64)", body)
    sizes = {"i32": 4, "f64": 8}
    ctype = {ct.c_int32: "i32", ct.c_double: "f64"}
    want = []
    for n, t in capi.SpParams._fields_:
        want.append((n, ctype[t] if t in ctype else f"[i32; {t._length_}]"))
    assert rust == want
    assert sum(sizes.get(t, 4 * int(t[6:-1]) if t.startswith("[") else 0) for _, t in rust) + 4 == ct.sizeof(capi.SpParams)
