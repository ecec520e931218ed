"""The benchmark workloads at their own step counts and sizes.

* Navier-Stokes pulse + vortex (uniform flow M=0.5, 1e-3 p0 pulse of
  half-width 4, isentropic vortex 200 points upstream, 40-point buffers,
  sigma=0.2): 1000 FP64 steps at 512^2 against the oracle's 1000 steps
  (relative L2 <= 1e-12 per field), and 1000 steps at the full 4096^2 in
  the FP32 perturbation form against FP64 (relative p' = p - p_b L2 <= 1e-4).
* The batched sweep as configured: 256 Euler Gaussian-pulse cases with
  square grids from {128, 192, 256, 384, 512}^2 (seeded M, amplitude and
  half-width), advanced together; one case of every size checked against
  the oracle in FP64, every case's p' in the FP32 perturbation form
  against the FP64 batch.
"""
import numpy as np
import pytest

from conftest import rel_l2
from navierstokes_amd import config
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64, FP32P = 0, 1


def _solver(case, prec=FP64):
    from navierstokes_amd.solver import Solver
    print(f"[case] {case.name} {case.imax}x{case.jmax} nk={case.nk} precision={prec}", flush=True)
    return Solver(case, precision=prec)


def test_ns_pulse_vortex_1000_steps_fp64_vs_oracle():
    case = config.pulse_vortex_case(512, D=40)
    s = _solver(case)
    s.run(0, 1000)
    g = s.get_state()
    s.close()
    print("[oracle] 1000 steps of 512^2 (OpenMP)", flush=True)
    o = Oracle(case, parallel=True).run(case.U0, 0, 1000)
    errs = [rel_l2(g[..., k], o[..., k]) for k in range(case.nk)]
    chg = [rel_l2(g[..., k] - case.U0[..., k], o[..., k] - case.U0[..., k]) for k in range(case.nk)]
    pp = rel_l2(g[..., 3] - case.Ub[..., 3], o[..., 3] - case.Ub[..., 3])
    print(f"relative L2 per field {errs}; of the change since step 0 {chg}; p' {pp:.3e}", flush=True)
    # the signal is still on the grid: the vortex has moved 167 points, the
    # pulse has spread but not left
    assert np.abs(o[..., 2]).max() > 1e-2
    assert max(errs) <= 1e-12
    assert pp <= 1e-9


def test_ns_pulse_vortex_4096_1000_steps_fp32_perturbation():
    case = config.pulse_vortex_case(4096, D=40)
    s64 = _solver(case, FP64)
    s64.run(0, 1000)
    b = s64.get_state()[..., 3] - case.Ub[..., 3]
    s64.close()
    s32 = _solver(case, FP32P)
    s32.run(0, 1000)
    a = s32.get_state()[..., 3] - case.Ub[..., 3]
    s32.close()
    err = rel_l2(a, b)
    print(f"p' relative L2 (NS pulse+vortex 4096^2, 1000 steps, FP32-pert vs FP64): {err:.3e}", flush=True)
    assert np.abs(b).max() > 1e-5
    assert err <= 1e-4


def test_batched_sweep_256_cases():
    from navierstokes_amd.solver import BatchSolver
    cases = config.batch_sweep_cases(256, seed=2024)
    npts = sum(c.npoints for c in cases)
    sizes = sorted({c.imax for c in cases})
    print(f"[case] batched sweep: {len(cases)} cases, sizes {sizes}, {npts / 1e6:.2f} Mpts", flush=True)
    assert sizes == [128, 192, 256, 384, 512]
    nsteps = 6
    b64 = BatchSolver(cases, precision=FP64)
    b64.run(0, nsteps)
    first = {}
    for i, c in enumerate(cases):
        first.setdefault(c.imax, i)
    for n, i in sorted(first.items()):
        c = cases[i]
        g = b64.get_state(i)
        o = Oracle(c).run(c.U0, 0, nsteps)
        err = max(rel_l2(g[..., k], o[..., k]) for k in range(c.nk))
        errc = rel_l2(g[..., 3] - c.U0[..., 3], o[..., 3] - c.U0[..., 3])
        print(f"case {i} ({c.name}): FP64 batch vs oracle {err:.3e} (of dp {errc:.3e}), "
              f"bitwise {np.array_equal(g, o)}", flush=True)
        assert err <= 1e-13 and errc <= 1e-11
    b32 = BatchSolver(cases, precision=FP32P)
    b32.run(0, nsteps)
    worst = 0.0
    for i, c in enumerate(cases):
        a, r = b32.get_state(i), b64.get_state(i)
        err = rel_l2(a[..., 3] - c.Ub[..., 3], r[..., 3] - c.Ub[..., 3])
        worst = max(worst, err)
    print(f"FP32-pert batch p' vs FP64 batch, worst of {len(cases)} cases: {worst:.3e}", flush=True)
    assert worst <= 1e-4
    b32.close()
    b64.close()
