"""Python mirror of the reference's hot-path interface over libnsamd.so.

`Solver` owns one device context (what `module global` holds in the
reference) and exposes the reference routines under their own names:

    rk_euler(iter)                 rk.f90:12      one LDDRK6 step
    run(iter0, nsteps)             euler.f90:123  the time loop
    rhs_euler(U, ct, iter)         rhs.f90:10
    deropm(U) / deropn(U)          diff.f90:6 / :127 (+ metrics)
    molecular_mi_k(T)              coefficients.f90:27-59
    rk_steep(a, b, U, dudt, w)     rk.f90:153
    bufferxy(U)                    rk.f90:77
    lddfilterx(U) / lddfiltery(U)  filtering.f90:9 / :98
    ccpml(U)                       cc.f90:6
    acousticfont(ct, iter)         rk.f90:171

Reverse mode (no reference counterpart; 2-D Euler): adj_set_arc /
adj_forward / adj_backward and the transposed operators vjp_*.

Host arrays follow the reference layout U(1:nk,1:imax,1:jmax), i.e. numpy
shape (jmax, imax, nk), C-contiguous, float64.  Errors raise NSError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import Case


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def make_params(case: Case, device: int = 0, precision: int = _lib.NS_PREC_FP64) -> _lib.NSParams:
    """ns_params of a case (the reference's `module global` scalars)."""
    lib = _lib.load()
    p = _lib.NSParams()
    _lib.check(lib.ns_default_params(C.byref(p)))
    tc = case.transport()
    p.imax, p.jmax, p.nk, p.D = case.imax, case.jmax, case.nk, case.D
    p.dm, p.dn, p.dt, p.sigmaf = case.dm, case.dn, case.dt, case.sigmaf
    p.gamma, p.Re, p.Pr, p.Pe = case.gamma, case.Re, case.Pr, case.Pe
    p.alpha_cond, p.Tref = case.alpha_cond, case.Tref
    for key in ("A_mi", "B_mi", "C_mi", "D_mi", "MI0", "A_k", "B_k", "C_k", "D_k", "K0"):
        setattr(p, key, tc[key])
    p.amplitude, p.wf, p.r0_acous = case.amplitude, case.wf, case.r0_acous
    p.m_pt_acous, p.n_pt_acous = case.m_pt_acous, case.n_pt_acous
    p.precision = precision
    flags = 0
    if case.inviscid:
        flags |= _lib.NS_FLAG_INVISCID
    if case.metric_everywhere:
        flags |= _lib.NS_FLAG_METRIC_EVERYWHERE
    if getattr(case, "periodic_x", False):
        flags |= _lib.NS_FLAG_PERIODIC_X
    p.flags = flags
    p.device = device
    return p


def _profiles(case: Case):
    c = case
    return [_f64(x) for x in (c.meshx, c.meshy, c.sigmax, c.sigmay, c.m, c.n, c.rhob, c.psib)]


class Solver:
    def __init__(self, case: Case, device: int = 0, precision: int = _lib.NS_PREC_FP64):
        self.case = case
        self.lib = _lib.load()
        p = make_params(case, device, precision)
        self.params = p
        ctx = C.c_void_p()
        _lib.check(self.lib.ns_create(C.byref(p), C.byref(ctx)))
        self.ctx = ctx
        self.shape = (case.jmax, case.imax, case.nk)
        c = case
        self._prof = _profiles(c)
        _lib.check(self.lib.ns_set_profiles(self.ctx, *[_lib.dptr(x) for x in self._prof]))
        self.set_base(c.Ub)
        self.put_state(c.U0)

    # lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.ns_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # state --------------------------------------------------------------------
    def set_base(self, Ub):
        Ub = _f64(Ub, self.shape)
        _lib.check(self.lib.ns_set_base(self.ctx, _lib.dptr(Ub)))

    def put_state(self, U):
        U = _f64(U, self.shape)
        _lib.check(self.lib.ns_put_state(self.ctx, _lib.dptr(U)))

    def get_state(self) -> np.ndarray:
        U = np.empty(self.shape)
        _lib.check(self.lib.ns_get_state(self.ctx, _lib.dptr(U)))
        return U

    def probe(self, points, noalloc: bool = False) -> np.ndarray:
        """Values at 1-based (i, j) points: shape (npts, nk) (euler.f90:156-161).
        noalloc: stage through the library's fixed device array instead of
        allocating (at most NS_PROBE_NOALLOC_MAX points)."""
        ij = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 2))
        out = np.empty((ij.shape[0], self.case.nk))
        fn = self.lib.ns_probe_noalloc if noalloc else self.lib.ns_probe
        _lib.check(fn(self.ctx, ij.shape[0], ij.ctypes.data_as(C.POINTER(C.c_int32)), _lib.dptr(out)))
        return out

    def vorticity(self) -> np.ndarray:
        """dv/dm - du/dn of the current state, shape (jmax, imax) (outt.f90:49-55)."""
        out = np.empty((self.case.jmax, self.case.imax))
        _lib.check(self.lib.ns_vorticity(self.ctx, _lib.dptr(out)))
        return out

    def station_energy(self, cols) -> np.ndarray:
        """Kinetic energy of columns `cols` (1-based) integrated over y (outt.f90:92-110)."""
        c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32).ravel())
        out = np.empty(c.shape[0])
        _lib.check(self.lib.ns_station_energy(self.ctx, c.shape[0], c.ctypes.data_as(C.POINTER(C.c_int32)),
                                              _lib.dptr(out)))
        return out

    def row_average(self, k: int, i0: int = 1, i1: int | None = None) -> np.ndarray:
        """Mean of component k over columns i0..i1 (1-based, inclusive) per row."""
        i1 = self.case.imax if i1 is None else i1
        out = np.empty(self.case.jmax)
        _lib.check(self.lib.ns_row_average(self.ctx, int(k), int(i0), int(i1), _lib.dptr(out)))
        return out

    def check_finite(self):
        _lib.check(self.lib.ns_check_finite(self.ctx))

    def synchronize(self):
        _lib.check(self.lib.ns_synchronize(self.ctx))

    # time advance -------------------------------------------------------------
    def rk_euler(self, iter: int):
        _lib.check(self.lib.ns_rk_euler(self.ctx, int(iter)))

    def run(self, iter0: int, nsteps: int):
        _lib.check(self.lib.ns_rk_steps(self.ctx, int(iter0), int(nsteps)))

    # operator mirrors ---------------------------------------------------------
    def deropm(self, U) -> np.ndarray:
        U = _f64(U)
        nk = U.shape[-1] if U.ndim == 3 else 1
        out = np.empty_like(U)
        _lib.check(self.lib.ns_deropm(self.ctx, nk, _lib.dptr(out), _lib.dptr(U)))
        return out

    def deropn(self, U) -> np.ndarray:
        U = _f64(U)
        nk = U.shape[-1] if U.ndim == 3 else 1
        out = np.empty_like(U)
        _lib.check(self.lib.ns_deropn(self.ctx, nk, _lib.dptr(out), _lib.dptr(U)))
        return out

    def molecular_mi_k(self, T):
        T = _f64(T, (self.case.jmax, self.case.imax))
        mi, kk = np.empty_like(T), np.empty_like(T)
        _lib.check(self.lib.ns_transport(self.ctx, _lib.dptr(T), _lib.dptr(mi), _lib.dptr(kk)))
        return mi, kk

    def rhs_euler(self, U, ct: float = 0.0, iter: int = 0) -> np.ndarray:
        U = _f64(U, self.shape)
        out = np.empty_like(U)
        _lib.check(self.lib.ns_rhs(self.ctx, _lib.dptr(out), _lib.dptr(U), float(ct), int(iter)))
        return out

    def rk_steep(self, alphat: float, betat: float, U, dudt, w):
        U = _f64(U, self.shape).copy()
        w = _f64(w, self.shape).copy()
        dudt = _f64(dudt, self.shape)
        _lib.check(self.lib.ns_rk_steep(self.ctx, float(alphat), float(betat), _lib.dptr(U),
                                        _lib.dptr(dudt), _lib.dptr(w)))
        return U, w

    def _inplace(self, fn, U):
        U = _f64(U, self.shape).copy()
        _lib.check(fn(self.ctx, _lib.dptr(U)))
        return U

    def bufferxy(self, U):
        return self._inplace(self.lib.ns_bufferxy, U)

    def lddfilterx(self, U):
        return self._inplace(self.lib.ns_filter_x, U)

    def lddfiltery(self, U):
        return self._inplace(self.lib.ns_filter_y, U)

    def ccpml(self, U):
        return self._inplace(self.lib.ns_ccpml, U)

    def acousticfont(self, ct: float, iter: int) -> np.ndarray:
        s = np.empty((self.case.jmax, self.case.imax))
        _lib.check(self.lib.ns_acousticfont(self.ctx, float(ct), int(iter), _lib.dptr(s)))
        return s

    # device-resident state (torch tensors of shape (nk, jmax, imax), float64) --
    def put_state_dev(self, U):
        self._check_dev(U)
        _lib.check(self.lib.ns_put_state_dev(self.ctx, C.c_void_p(U.data_ptr())))

    def get_state_dev(self, out):
        self._check_dev(out)
        _lib.check(self.lib.ns_get_state_dev(self.ctx, C.c_void_p(out.data_ptr())))

    def _check_dev(self, t):
        c = self.case
        if (tuple(t.shape) != (c.nk, c.jmax, c.imax) or str(t.dtype) != "torch.float64" or not t.is_cuda
                or not t.is_contiguous()):
            raise ValueError(f"expected a contiguous float64 device tensor of shape {(c.nk, c.jmax, c.imax)}")

    # reverse mode -------------------------------------------------------------
    def adj_set_arc(self, points):
        """Observer points, 1-based (i, j), each once."""
        ij = np.ascontiguousarray(np.asarray(points, dtype=np.int32).reshape(-1, 2))
        _lib.check(self.lib.ns_adj_set_arc(self.ctx, ij.shape[0], ij.ctypes.data_as(C.POINTER(C.c_int32))))

    def adj_forward(self, iter0: int, nsteps: int, every: int) -> float:
        """Runs nsteps from the current state keeping checkpoints; returns J.
        Each call replaces the checkpoints an earlier one kept (the backward
        always differentiates the latest forward): `adj_generation` counts
        the calls."""
        J = C.c_double()
        _lib.check(self.lib.ns_adj_forward(self.ctx, int(iter0), int(nsteps), int(every), C.byref(J)))
        self.adj_generation = getattr(self, "adj_generation", 0) + 1
        return J.value

    def adj_set_keep(self, keep: bool):
        """Whether the backward may keep a segment's stage inputs (default)
        or recomputes them per step; both give the same bits."""
        _lib.check(self.lib.ns_adj_set_keep(self.ctx, int(bool(keep))))

    def adj_backward(self) -> np.ndarray:
        out = np.empty(self.shape)
        _lib.check(self.lib.ns_adj_backward(self.ctx, _lib.dptr(out)))
        return out

    def adj_backward_dev(self, out):
        self._check_dev(out)
        _lib.check(self.lib.ns_adj_backward_dev(self.ctx, C.c_void_p(out.data_ptr())))

    def vjp_rhs(self, U, lam_f) -> np.ndarray:
        U = _f64(U, self.shape)
        lam_f = _f64(lam_f, self.shape)
        out = np.empty(self.shape)
        _lib.check(self.lib.ns_vjp_rhs(self.ctx, _lib.dptr(U), _lib.dptr(lam_f), _lib.dptr(out)))
        return out

    def vjp_filter_x(self, lam):
        return self._inplace(self.lib.ns_vjp_filter_x, lam)

    def vjp_filter_y(self, lam):
        return self._inplace(self.lib.ns_vjp_filter_y, lam)

    def vjp_ccpml(self, lam):
        return self._inplace(self.lib.ns_vjp_ccpml, lam)

    def vjp_step(self, iter: int, U, lam) -> np.ndarray:
        U = _f64(U, self.shape)
        lam = _f64(lam, self.shape).copy()
        _lib.check(self.lib.ns_vjp_step(self.ctx, int(iter), _lib.dptr(U), _lib.dptr(lam)))
        return lam

    # tuning / measurement -----------------------------------------------------
    def autotune(self):
        """Pick the fastest stage-kernel block size; returns (choice, ms[3])."""
        v = C.c_int32()
        ms = np.zeros(3)
        _lib.check(self.lib.ns_autotune(self.ctx, C.byref(v), _lib.dptr(ms)))
        return v.value, ms

    def set_variant(self, variant: int):
        """Use stage variant 0..2 (see autotune) without timing."""
        _lib.check(self.lib.ns_set_variant(self.ctx, int(variant)))

    def set_filter_mode(self, mode: int):
        """0: the x and y filter sweeps as two kernels; 1: one fused sweep."""
        _lib.check(self.lib.ns_set_filter_mode(self.ctx, int(mode)))

    def filter_mode(self):
        """(mode, [ms of the two sweeps, ms of the fused sweep]) (-1: not tuned)."""
        m = C.c_int32()
        ms = np.zeros(2)
        _lib.check(self.lib.ns_get_filter_mode(self.ctx, C.byref(m), _lib.dptr(ms)))
        return m.value, ms

    def enable_timing(self, on: bool = True):
        _lib.check(self.lib.ns_enable_timing(self.ctx, int(on)))

    def stage_interior_points(self) -> int:
        """Output points of the stage's interior kernel in the last timed stage."""
        n = C.c_int64()
        _lib.check(self.lib.ns_stage_interior(self.ctx, C.byref(n)))
        return n.value

    def kernel_time(self, name: str):
        avg = C.c_double()
        cnt = C.c_int32()
        _lib.check(self.lib.ns_kernel_time(self.ctx, name.encode(), C.byref(avg), C.byref(cnt)))
        return avg.value, cnt.value


class BatchSolver:
    """Many independent cases (any grid sizes) advanced together.

    The reference runs one case per process (euler.f90:123-170 is the whole
    time loop); here every kernel class of a step runs as one launch over all
    cases and a captured HIP graph replays the batch step.  All cases share
    nk, precision and flags, and run without the acoustic source.
    """

    def __init__(self, cases, device: int = 0, precision: int = _lib.NS_PREC_FP64):
        self.cases = list(cases)
        self.lib = _lib.load()
        n = len(self.cases)
        arr = (_lib.NSParams * n)(*[make_params(c, device, precision) for c in self.cases])
        b = C.c_void_p()
        _lib.check(self.lib.ns_batch_create(arr, n, C.byref(b)))
        self.b = b
        for i, c in enumerate(self.cases):
            shape = (c.jmax, c.imax, c.nk)
            prof = _profiles(c)
            Ub, U0 = _f64(c.Ub, shape), _f64(c.U0, shape)
            _lib.check(self.lib.ns_batch_set_case(self.b, i, *[_lib.dptr(x) for x in prof],
                                                  _lib.dptr(Ub), _lib.dptr(U0)))

    def close(self):
        if getattr(self, "b", None):
            self.lib.ns_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, iter0: int, nsteps: int):
        _lib.check(self.lib.ns_batch_steps(self.b, int(iter0), int(nsteps)))

    def get_state(self, c: int) -> np.ndarray:
        cs = self.cases[c]
        U = np.empty((cs.jmax, cs.imax, cs.nk))
        _lib.check(self.lib.ns_batch_get_state(self.b, int(c), _lib.dptr(U)))
        return U

    def autotune(self):
        """Time the stage variants (0: tiles, 1: marching with fast zones,
        2: marching on plain general regions) and keep the fastest; returns
        (choice, ms[3]), -1 for a variant that cannot run these cases."""
        v = C.c_int32()
        ms = np.zeros(3)
        _lib.check(self.lib.ns_batch_autotune(self.b, C.byref(v), _lib.dptr(ms)))
        return v.value, ms

    def set_variant(self, variant: int):
        """Use stage variant 0..2 (see autotune)."""
        _lib.check(self.lib.ns_batch_set_variant(self.b, int(variant)))
