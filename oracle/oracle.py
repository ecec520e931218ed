"""ctypes wrapper of the CPU oracle (ns_oracle.c) and of the partial reference
build (oracle/_ref/libnsref.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_SERIAL = HERE / "libns_oracle.so"
LIB_OMP = HERE / "libns_oracle_omp.so"
LIB_REF = HERE / "_ref" / "libnsref.so"
REFERENCE_TREE = Path("/root/reference")

_DP = C.POINTER(C.c_double)


class NSOParams(C.Structure):
    _fields_ = [
        ("imax", C.c_int), ("jmax", C.c_int), ("nk", C.c_int), ("D", C.c_int),
        ("dm", C.c_double), ("dn", C.c_double), ("dt", C.c_double), ("sigmaf", C.c_double),
        ("gamma", C.c_double), ("Re", C.c_double), ("Pr", C.c_double), ("Pe", C.c_double),
        ("alpha_cond", C.c_double), ("Tref", C.c_double),
        ("A_mi", C.c_double), ("B_mi", C.c_double), ("C_mi", C.c_double), ("D_mi", C.c_double),
        ("MI0", C.c_double),
        ("A_k", C.c_double), ("B_k", C.c_double), ("C_k", C.c_double), ("D_k", C.c_double),
        ("K0", C.c_double),
        ("amplitude", C.c_double), ("wf", C.c_double), ("r0_acous", C.c_double),
        ("m_pt_acous", C.c_double), ("n_pt_acous", C.c_double),
        ("meshx", _DP), ("meshy", _DP), ("sigmax", _DP), ("sigmay", _DP), ("m", _DP),
        ("n", _DP), ("rhob", _DP), ("psib", _DP), ("Ub", _DP),
        ("metric_everywhere", C.c_int),
        ("periodic_x", C.c_int),
    ]


RHS_FN = C.CFUNCTYPE(None, C.POINTER(NSOParams), _DP, _DP, C.c_int, C.c_double, C.c_void_p)


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and the reference's compilable files when the
    reference tree is present)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if with_ref and REFERENCE_TREE.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


_libs: dict = {}


def _load(path: Path) -> C.CDLL:
    if path not in _libs:
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `make -C {HERE}`")
        lib = C.CDLL(str(path))
        for name in ("nso_deropm", "nso_deropn"):
            getattr(lib, name).argtypes = [C.POINTER(NSOParams), C.c_int, _DP, _DP]
        lib.nso_transport.argtypes = [C.POINTER(NSOParams), _DP, _DP, _DP]
        lib.nso_rhs.argtypes = [C.POINTER(NSOParams), _DP, _DP, _DP]
        lib.nso_rk_steep.argtypes = [C.POINTER(NSOParams), C.c_double, C.c_double, _DP, _DP, _DP]
        for name in ("nso_bufferxy", "nso_filterx", "nso_filtery", "nso_ccpml"):
            getattr(lib, name).argtypes = [C.POINTER(NSOParams), _DP]
        lib.nso_acousticfont.argtypes = [C.POINTER(NSOParams), C.c_double, C.c_int, _DP]
        lib.nso_rk_euler.argtypes = [C.POINTER(NSOParams), _DP, C.c_int]
        lib.nso_rk_euler_cb.argtypes = [C.POINTER(NSOParams), _DP, C.c_int, RHS_FN, C.c_void_p]
        lib.nso_run.argtypes = [C.POINTER(NSOParams), _DP, C.c_int, C.c_int]
        lib.nso_lddrk_coeffs.argtypes = [_DP, _DP, _DP]
        lib.nso_get_coefficients.argtypes = [C.POINTER(NSOParams)]
        lib.nso_num_threads.restype = C.c_int
        _libs[path] = lib
    return _libs[path]


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """FP64 restatement of the reference hot path for one Case."""

    def __init__(self, case, parallel: bool = False):
        self.case = case
        self.lib = _load(LIB_OMP if parallel else LIB_SERIAL)
        c = case
        tc = c.transport()
        self._keep = [_c(x) for x in (c.meshx, c.meshy, c.sigmax, c.sigmay, c.m, c.n, c.rhob, c.psib, c.Ub)]
        p = NSOParams()
        p.imax, p.jmax, p.nk, p.D = c.imax, c.jmax, c.nk, c.D
        p.dm, p.dn, p.dt, p.sigmaf = c.dm, c.dn, c.dt, c.sigmaf
        p.gamma, p.Re, p.Pr, p.Pe, p.alpha_cond, p.Tref = c.gamma, c.Re, c.Pr, c.Pe, c.alpha_cond, c.Tref
        for key in ("A_mi", "B_mi", "C_mi", "D_mi", "MI0", "A_k", "B_k", "C_k", "D_k", "K0"):
            setattr(p, key, tc[key])
        p.amplitude, p.wf, p.r0_acous = c.amplitude, c.wf, c.r0_acous
        p.m_pt_acous, p.n_pt_acous = c.m_pt_acous, c.n_pt_acous
        (p.meshx, p.meshy, p.sigmax, p.sigmay, p.m, p.n, p.rhob, p.psib, p.Ub) = [
            x.ctypes.data_as(_DP) for x in self._keep]
        p.metric_everywhere = int(bool(getattr(c, "metric_everywhere", False)))
        p.periodic_x = int(bool(getattr(c, "periodic_x", False)))
        self.p = p
        self.shape = (c.jmax, c.imax, c.nk)

    def threads(self) -> int:
        return self.lib.nso_num_threads()

    def deropm(self, U):
        U = _c(U)
        nk = U.shape[-1] if U.ndim == 3 else 1
        out = np.empty_like(U)
        self.lib.nso_deropm(C.byref(self.p), nk, out.ctypes.data_as(_DP), U.ctypes.data_as(_DP))
        return out

    def deropn(self, U):
        U = _c(U)
        nk = U.shape[-1] if U.ndim == 3 else 1
        out = np.empty_like(U)
        self.lib.nso_deropn(C.byref(self.p), nk, out.ctypes.data_as(_DP), U.ctypes.data_as(_DP))
        return out

    def transport(self, T):
        T = _c(T)
        mi, kk = np.empty_like(T), np.empty_like(T)
        self.lib.nso_transport(C.byref(self.p), T.ctypes.data_as(_DP), mi.ctypes.data_as(_DP),
                               kk.ctypes.data_as(_DP))
        return mi, kk

    def rhs(self, U, s_acou=None):
        U = _c(U)
        out = np.empty_like(U)
        s = None if s_acou is None else _c(s_acou).ctypes.data_as(_DP)
        self.lib.nso_rhs(C.byref(self.p), out.ctypes.data_as(_DP), U.ctypes.data_as(_DP), s)
        return out

    def rk_steep(self, alphat, betat, U, dudt, w):
        U, w, dudt = _c(U).copy(), _c(w).copy(), _c(dudt)
        self.lib.nso_rk_steep(C.byref(self.p), alphat, betat, U.ctypes.data_as(_DP),
                              dudt.ctypes.data_as(_DP), w.ctypes.data_as(_DP))
        return U, w

    def _inplace(self, fn, U):
        U = _c(U).copy()
        fn(C.byref(self.p), U.ctypes.data_as(_DP))
        return U

    def bufferxy(self, U):
        return self._inplace(self.lib.nso_bufferxy, U)

    def filterx(self, U):
        return self._inplace(self.lib.nso_filterx, U)

    def filtery(self, U):
        return self._inplace(self.lib.nso_filtery, U)

    def ccpml(self, U):
        return self._inplace(self.lib.nso_ccpml, U)

    def acousticfont(self, ct, it):
        s = np.empty((self.case.jmax, self.case.imax))
        self.lib.nso_acousticfont(C.byref(self.p), ct, it, s.ctypes.data_as(_DP))
        return s

    def rk_euler(self, U, it):
        U = _c(U).copy()
        self.lib.nso_rk_euler(C.byref(self.p), U.ctypes.data_as(_DP), it)
        return U

    def rk_euler_cb(self, U, it, fn):
        """Stage loop with a Python RHS: fn(U, irk, t) -> dudt (same shape)."""
        shape = self.shape

        def cb(pp, dudt, Uin, irk, t, user):
            n = shape[0] * shape[1] * shape[2]
            u = np.ctypeslib.as_array(Uin, shape=(n,)).reshape(shape)
            d = np.ctypeslib.as_array(dudt, shape=(n,)).reshape(shape)
            d[...] = fn(u.copy(), irk, t)

        cfn = RHS_FN(cb)
        U = _c(U).copy()
        self.lib.nso_rk_euler_cb(C.byref(self.p), U.ctypes.data_as(_DP), it, cfn, None)
        return U

    def run(self, U, iter0, nsteps):
        U = _c(U).copy()
        self.lib.nso_run(C.byref(self.p), U.ctypes.data_as(_DP), iter0, nsteps)
        return U

    def lddrk_coeffs(self):
        a, b, c = np.empty(6), np.empty(6), np.empty(6)
        self.lib.nso_lddrk_coeffs(a.ctypes.data_as(_DP), b.ctypes.data_as(_DP), c.ctypes.data_as(_DP))
        return a, b, c


class Reference:
    """The reference's own compilable routines (diffall.f90, coefficients.f90,
    cc.f90) built from the reference tree into oracle/_ref/libnsref.so."""

    def __init__(self, imax, jmax, dm=1.0, dn=1.0, gamma=1.4, Tref=273.15):
        lib = C.CDLL(str(LIB_REF))
        lib.nsref_setup.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        lib.nsref_get_coefficients.argtypes = [_DP]
        lib.nsref_transport.argtypes = [_DP, _DP, _DP]
        lib.nsref_ccpml.argtypes = [_DP, C.c_int, _DP, _DP]
        lib.nsref_deropm.argtypes = [_DP, _DP, C.c_int]
        lib.nsref_deropn.argtypes = [_DP, _DP, C.c_int]
        self.lib = lib
        self.imax, self.jmax = imax, jmax
        self._args = (imax, jmax, dm, dn, gamma, Tref)
        self._setup()

    def _setup(self):
        # the reference keeps its state in module globals: re-apply per call
        self.lib.nsref_setup(*self._args)

    @staticmethod
    def available() -> bool:
        return LIB_REF.exists()

    def get_coefficients(self):
        self._setup()
        out = np.empty(10)
        self.lib.nsref_get_coefficients(out.ctypes.data_as(_DP))
        keys = ["A_mi", "B_mi", "C_mi", "D_mi", "A_k", "B_k", "C_k", "D_k", "MI0", "K0"]
        return dict(zip(keys, out))

    def transport(self, T):
        self._setup()
        T = _c(T)
        mi, kk = np.empty_like(T), np.empty_like(T)
        self.lib.nsref_transport(T.ctypes.data_as(_DP), mi.ctypes.data_as(_DP), kk.ctypes.data_as(_DP))
        return mi, kk

    def ccpml(self, U, rhob, psib):
        self._setup()
        U = _c(U).copy()
        assert U.shape[-1] == 5, "cc.f90 writes component 5: nk must be 5"
        rb, pb = _c(rhob), _c(psib)
        self.lib.nsref_ccpml(U.ctypes.data_as(_DP), 5, rb.ctypes.data_as(_DP), pb.ctypes.data_as(_DP))
        return U

    def deropm(self, U):
        self._setup()
        U = _c(U).copy()
        nk = U.shape[-1] if U.ndim == 3 else 1
        out = np.zeros_like(U)
        self.lib.nsref_deropm(out.ctypes.data_as(_DP), U.ctypes.data_as(_DP), nk)
        return out

    def deropn(self, U):
        self._setup()
        U = _c(U).copy()
        nk = U.shape[-1] if U.ndim == 3 else 1
        out = np.zeros_like(U)
        self.lib.nsref_deropn(out.ctypes.data_as(_DP), U.ctypes.data_as(_DP), nk)
        return out
