"""C ABI checks that need no GPU: the library loads, exports every entry point
that include/ns_api.h declares, and validates parameters before it touches a
device."""
import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from navierstokes_amd import _lib, config

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "ns_api.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ns_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True)
    exported = set(re.findall(r"\bT (ns_\w+)", out.stdout))
    assert set(names) <= exported


def test_struct_layout_matches_header(tmp_path):
    """ctypes NSParams has the field offsets gcc gives the header's struct."""
    fields = [f for f, _ in _lib.NSParams._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(ns_params, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ns_api.h"\n'
                   'int main(void){printf("sizeof %zu\\n", sizeof(ns_params));' + body + 'return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    assert int(got["sizeof"]) == C.sizeof(_lib.NSParams)
    for f in fields:
        assert int(got[f]) == getattr(_lib.NSParams, f).offset, f


def test_default_params_are_reference_values():
    lib = _lib.load()
    p = _lib.NSParams()
    assert lib.ns_default_params(C.byref(p)) == 0
    tc = config.transport_coefficients(273.15)
    assert p.MI0 == tc["MI0"] and p.K0 == tc["K0"]
    assert (p.imax, p.jmax, p.nk, p.D) == (521, 481, 5, 90)   # dados.in
    assert p.gamma == 1.4 and p.Pe == 1000.0 and p.alpha_cond == 0.5
    assert lib.ns_api_version() == 1


def _params(**kw):
    p = _lib.NSParams()
    _lib.load().ns_default_params(C.byref(p))
    p.imax, p.jmax, p.nk, p.D = 64, 48, 4, 8
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("field,value,msg", [
    ("imax", 8, "imax and jmax"),
    ("nk", 3, "nk must be"),
    ("D", 40, "buffer width"),
    ("dt", 0.0, "positive"),
    ("sigmaf", 1.5, "sigmaf"),
    ("precision", 7, "precision"),
])
def test_create_rejects_bad_parameters(field, value, msg):
    lib = _lib.load()
    ctx = C.c_void_p()
    rc = lib.ns_create(C.byref(_params(**{field: value})), C.byref(ctx))
    assert rc == _lib.NS_ERR_PARAM
    assert msg in lib.ns_last_error().decode()
    assert not ctx.value


def test_null_arguments_fail_cleanly():
    lib = _lib.load()
    assert lib.ns_create(None, None) == _lib.NS_ERR_PARAM
    assert lib.ns_rk_euler(None, 0) == _lib.NS_ERR_PARAM
    assert lib.ns_destroy(None) == 0


def test_no_device_reports_hip_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    ctx = C.c_void_p()
    rc = lib.ns_create(C.byref(_params()), C.byref(ctx))
    assert rc == _lib.NS_ERR_HIP
    assert "hipSetDevice" in lib.ns_last_error().decode()


def test_solver_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="only compute path"):
        _lib.load()
