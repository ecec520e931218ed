"""Host-side accuracy of the FP64 log and exp of the transport coefficients.

transport_math_check.cpp includes navierstokes_amd/csrc/ns_device.h and
measures ns_log / ns_exp (the functions the FP64 kernels evaluate mu(T) and
k(T) with, coefficients.f90:35-42, :50-58) against the C library in long
double on 2 million random arguments over the transport's range and on the
edges of the argument reductions: within 1.5 ulp (log) and 1 ulp (exp), and
insensitive (<= 1 ulp) to the reciprocal estimate the Newton steps start
from.  A dropped polynomial term or a wrong ln 2 split fails it.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    return None


@pytest.mark.skipif(_hipcc() is None, reason="hipcc (host compiler for the HIP headers) not available")
def test_transport_log_exp_accuracy(tmp_path):
    exe = tmp_path / "transport_math_check"
    subprocess.run([_hipcc(), "-std=c++17", "-O2", "-I", str(ROOT / "navierstokes_amd" / "csrc"), "-I",
                    str(ROOT / "include"), str(ROOT / "tests" / "transport_math_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    f = res.stdout.split()
    assert float(f[1]) <= 1.5 and float(f[3]) <= 1.0
