"""ns_probe_noalloc (no reference counterpart): ns_probe staged through the
library's fixed per-device array instead of a device allocation.  Called
through the C ABI and checked against ns_probe bitwise, twice in a row with
the second call smaller, in both precision modes; sizes beyond its maximum
are refused."""
import ctypes as C

import numpy as np
import pytest

from navierstokes_amd import config

pytestmark = pytest.mark.gpu


def _points(case, n, seed):
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.integers(1, case.imax + 1, n), rng.integers(1, case.jmax + 1, n)], axis=1)
    pts[:4] = [(1, 1), (case.imax, 1), (1, case.jmax), (case.imax, case.jmax)]
    return np.ascontiguousarray(pts.astype(np.int32))


def _call(lib, ctx, fn, pts, nk):
    out = np.full((pts.shape[0], nk), np.nan)
    rc = getattr(lib, fn)(ctx, pts.shape[0], pts.ctypes.data_as(C.POINTER(C.c_int32)),
                          out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


@pytest.mark.parametrize("prec", [0, 1])
def test_probe_noalloc_matches_probe(prec):
    from navierstokes_amd import _lib
    from navierstokes_amd.solver import Solver
    case = config.random_case(96, 72, nk=5, D=8, seed=41)
    print(f"[case] ns_probe_noalloc {case.name} precision={prec}", flush=True)
    s = Solver(case, precision=prec)
    s.run(0, 2)
    lib, ctx = _lib.load(), s.ctx
    for n, seed in ((3000, 1), (17, 2)):   # the second call smaller than the first
        pts = _points(case, n, seed)
        rc0, ref = _call(lib, ctx, "ns_probe", pts, case.nk)
        rc1, got = _call(lib, ctx, "ns_probe_noalloc", pts, case.nk)
        assert rc0 == 0 and rc1 == 0, _lib.load().ns_last_error()
        assert np.array_equal(got, ref), f"{n} points: max diff {np.nanmax(np.abs(got - ref)):.3e}"
    big = _points(case, _lib.NS_PROBE_NOALLOC_MAX + 1, 3)
    rc, _ = _call(lib, ctx, "ns_probe_noalloc", big, case.nk)
    assert rc == _lib.NS_ERR_PARAM
    assert b"NS_PROBE_NOALLOC_MAX" in lib.ns_last_error()
    rc, _ = _call(lib, ctx, "ns_probe_noalloc", np.zeros((0, 2), dtype=np.int32), case.nk)
    assert rc == 0
    s.close()
