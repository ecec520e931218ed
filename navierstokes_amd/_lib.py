"""ctypes binding of libnsamd.so (include/ns_api.h).

The library is the only compute path: if it is missing, importing the solver
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnsamd.so"

NS_OK = 0
NS_ERR_PARAM = -1
NS_ERR_HIP = -2
NS_ERR_NOMEM = -3
NS_ERR_STATE = -4
NS_ERR_NONFINITE = -5

NS_PREC_FP64 = 0
NS_PREC_FP32_PERT = 1

NS_FLAG_INVISCID = 1
NS_FLAG_METRIC_EVERYWHERE = 2
NS_FLAG_PERIODIC_X = 4

NS_PROBE_NOALLOC_MAX = 4096


class NSParams(C.Structure):
    _fields_ = [
        ("imax", C.c_int32), ("jmax", C.c_int32), ("nk", C.c_int32), ("D", C.c_int32),
        ("dm", C.c_double), ("dn", C.c_double), ("dt", C.c_double), ("sigmaf", C.c_double),
        ("gamma", C.c_double), ("Re", C.c_double), ("Pr", C.c_double), ("Pe", C.c_double),
        ("alpha_cond", C.c_double), ("Tref", C.c_double),
        ("A_mi", C.c_double), ("B_mi", C.c_double), ("C_mi", C.c_double), ("D_mi", C.c_double),
        ("MI0", C.c_double),
        ("A_k", C.c_double), ("B_k", C.c_double), ("C_k", C.c_double), ("D_k", C.c_double),
        ("K0", C.c_double),
        ("amplitude", C.c_double), ("wf", C.c_double), ("r0_acous", C.c_double),
        ("m_pt_acous", C.c_double), ("n_pt_acous", C.c_double),
        ("precision", C.c_int32), ("flags", C.c_uint32), ("device", C.c_int32),
        ("reserved", C.c_int32),
    ]


class NSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ns error {code}: {msg}")
        self.code = code


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_V = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "ns_api_version": (C.c_int32, []),
    "ns_last_error": (C.c_char_p, []),
    "ns_default_params": (C.c_int32, [C.POINTER(NSParams)]),
    "ns_create": (C.c_int32, [C.POINTER(NSParams), C.POINTER(_V)]),
    "ns_destroy": (C.c_int32, [_V]),
    "ns_set_stream": (C.c_int32, [_V, _V]),
    "ns_synchronize": (C.c_int32, [_V]),
    "ns_set_profiles": (C.c_int32, [_V, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP]),
    "ns_set_base": (C.c_int32, [_V, _DP]),
    "ns_put_state": (C.c_int32, [_V, _DP]),
    "ns_get_state": (C.c_int32, [_V, _DP]),
    "ns_probe": (C.c_int32, [_V, C.c_int32, _IP, _DP]),
    "ns_probe_noalloc": (C.c_int32, [_V, C.c_int32, _IP, _DP]),
    "ns_check_finite": (C.c_int32, [_V]),
    "ns_row_average": (C.c_int32, [_V, C.c_int32, C.c_int32, C.c_int32, _DP]),
    "ns_vorticity": (C.c_int32, [_V, _DP]),
    "ns_station_energy": (C.c_int32, [_V, C.c_int32, _IP, _DP]),
    "ns_rk_euler": (C.c_int32, [_V, C.c_int32]),
    "ns_rk_steps": (C.c_int32, [_V, C.c_int32, C.c_int32]),
    "ns_deropm": (C.c_int32, [_V, C.c_int32, _DP, _DP]),
    "ns_deropn": (C.c_int32, [_V, C.c_int32, _DP, _DP]),
    "ns_transport": (C.c_int32, [_V, _DP, _DP, _DP]),
    "ns_rhs": (C.c_int32, [_V, _DP, _DP, C.c_double, C.c_int32]),
    "ns_rk_steep": (C.c_int32, [_V, C.c_double, C.c_double, _DP, _DP, _DP]),
    "ns_bufferxy": (C.c_int32, [_V, _DP]),
    "ns_filter_x": (C.c_int32, [_V, _DP]),
    "ns_filter_y": (C.c_int32, [_V, _DP]),
    "ns_ccpml": (C.c_int32, [_V, _DP]),
    "ns_acousticfont": (C.c_int32, [_V, C.c_double, C.c_int32, _DP]),
    "ns_deriv": (C.c_int32, [_V, C.c_int32] + [_DP] * 12),
    "ns_kdt": (C.c_int32, [_V, _DP, _DP, _DP]),
    "ns_tau_ij": (C.c_int32, [_V, C.c_int32] + [_DP] * 6),
    "ns_eulernonlinear": (C.c_int32, [_V, C.c_int32] + [_DP] * 12),
    "ns_metrics_x": (C.c_int32, [_V, C.c_int32, _DP]),
    "ns_metrics_y": (C.c_int32, [_V, C.c_int32, _DP]),
    "ns_autotune": (C.c_int32, [_V, _IP, _DP]),
    "ns_set_variant": (C.c_int32, [_V, C.c_int32]),
    "ns_set_filter_mode": (C.c_int32, [_V, C.c_int32]),
    "ns_get_filter_mode": (C.c_int32, [_V, _IP, _DP]),
    "ns_stage_interior": (C.c_int32, [_V, C.POINTER(C.c_int64)]),
    "ns_enable_timing": (C.c_int32, [_V, C.c_int32]),
    "ns_kernel_time": (C.c_int32, [_V, C.c_char_p, _DP, _IP]),
    "ns_batch_create": (C.c_int32, [C.POINTER(NSParams), C.c_int32, C.POINTER(_V)]),
    "ns_batch_set_case": (C.c_int32, [_V, C.c_int32] + [_DP] * 10),
    "ns_batch_steps": (C.c_int32, [_V, C.c_int32, C.c_int32]),
    "ns_batch_get_state": (C.c_int32, [_V, C.c_int32, _DP]),
    "ns_batch_autotune": (C.c_int32, [_V, _IP, _DP]),
    "ns_batch_set_variant": (C.c_int32, [_V, C.c_int32]),
    "ns_batch_destroy": (C.c_int32, [_V]),
    "ns_put_state_dev": (C.c_int32, [_V, _V]),
    "ns_get_state_dev": (C.c_int32, [_V, _V]),
    "ns_adj_set_arc": (C.c_int32, [_V, C.c_int32, _IP]),
    "ns_adj_set_keep": (C.c_int32, [_V, C.c_int32]),
    "ns_adj_forward": (C.c_int32, [_V, C.c_int32, C.c_int32, C.c_int32, _DP]),
    "ns_adj_backward": (C.c_int32, [_V, _DP]),
    "ns_adj_backward_dev": (C.c_int32, [_V, _V]),
    "ns_vjp_rhs": (C.c_int32, [_V, _DP, _DP, _DP]),
    "ns_vjp_filter_x": (C.c_int32, [_V, _DP]),
    "ns_vjp_filter_y": (C.c_int32, [_V, _DP]),
    "ns_vjp_ccpml": (C.c_int32, [_V, _DP]),
    "ns_vjp_step": (C.c_int32, [_V, C.c_int32, _DP, _DP]),
}

_lib = None


def _torch_hip_runtime():
    """torch's own libamdhip64.so, if torch is installed (found without importing it)."""
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return None
    for loc in spec.submodule_search_locations:
        p = Path(loc) / "lib" / "libamdhip64.so"
        if p.exists():
            return p
    return None


def load() -> C.CDLL:
    """Load libnsamd.so (built by navierstokes_amd.build); raise if absent.

    One process must hold one HIP runtime, or device pointers (torch tensors)
    cannot cross between torch and the library.  torch loads its bundled
    libamdhip64.so by file name; the library needs libamdhip64.so.7, the same
    soname.  Preloading torch's copy first makes the library bind to it (soname
    match) and a later `import torch` reuse it (same file), in either import
    order.  Without torch the system ROCm runtime is used."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m navierstokes_amd.build` "
            "(the HIP library is the only compute path)")
    rt = _torch_hip_runtime()
    if rt is not None:
        C.CDLL(str(rt), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != NS_OK:
        msg = load().ns_last_error().decode(errors="replace")
        raise NSError(rc, msg)


def dptr(a):
    """double* of a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(_DP)
