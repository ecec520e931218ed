"""Build the gfx950 library libnsamd.so (HIP kernels + C ABI) in-tree.

The library has no torch types in its interface; Python reaches it through
ctypes (navierstokes_amd/_lib.py) and Fortran through ISO_C_BINDING
(navierstokes_amd/fortran/ns_shim.f90).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libnsamd.so"

SOURCES = ["ns_kernels.hip", "ns_adjoint.hip", "ns_api.hip"]
HEADERS = ["ns_device.h", "ns_kernels.h", "ns_adjoint.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed-FP32 instructions anywhere (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libnsamd.so)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")


def build(verbose: bool = False) -> Path:
    """Compile libnsamd.so for gfx950 (always recompiles).

    ns_kernels.hip and ns_adjoint.hip are compiled twice: the FP64 path with
    -ffp-contract=off (reference operation order, no fused multiply-adds) and
    the FP32 perturbation form with contraction."""
    cc = hipcc()
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC)]
    base = [f for f in HIPCC_FLAGS if f != "-shared"]
    bdir = PKG / "_build"
    bdir.mkdir(exist_ok=True)
    objs = []
    units = [
        ("ns_kernels.hip", "ns_kernels_f64.o", ["-DNS_KERNELS_F64", "-ffp-contract=off"]),
        ("ns_kernels.hip", "ns_kernels_f32.o", ["-DNS_KERNELS_F32"]),
        ("ns_adjoint.hip", "ns_adjoint_f64.o", ["-DNS_KERNELS_F64", "-ffp-contract=off"]),
        ("ns_adjoint.hip", "ns_adjoint_f32.o", ["-DNS_KERNELS_F32"]),
        ("ns_api.hip", "ns_api.o", []),
    ]
    # the units are independent: compile them concurrently
    from concurrent.futures import ThreadPoolExecutor
    cmds = []
    for src, obj, extra in units:
        out = bdir / obj
        cmds.append([cc, *base, *extra, *inc, "-c", str(CSRC / src), "-o", str(out)])
        objs.append(str(out))
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
