"""Discrete adjoint (reverse mode) of the 2-D Euler LDDRK6 step.

Checks, from the operator level up:
  * the transposed linear operators against the forward ones (dot products);
  * R'(U)^T and one step's transpose against central differences of the
    forward GPU path;
  * dJ/d(initial p) against central differences along random directions,
    against torch.autograd through the float64 torch restatement (itself
    pinned to the oracle), across checkpoint intervals, and through the
    torch.autograd.Function;
  * the FP32 perturbation form's gradient against FP64.
"""
import numpy as np
import pytest

from conftest import rel_l2
from navierstokes_amd import config

pytestmark = pytest.mark.gpu

FP64, FP32P = 0, 1


def _solver(case, prec=FP64):
    from navierstokes_amd.solver import Solver
    print(f"[case] {case.name} {case.imax}x{case.jmax} precision={prec}", flush=True)
    return Solver(case, precision=prec)


def _euler_case(nx=70, ny=60, seed=71):
    return config.random_case(nx, ny, nk=4, D=8, seed=seed, amp=1e-3, inviscid=True, mesh_random=True, Re=1e300)


def _dot(a, b):
    return float(np.sum(a * b))


def test_transposed_filters_and_ccpml():
    # 300 x 220: several look-ahead chunks per line (the transposed sweeps run
    # in chunks of 64)
    case = _euler_case(300, 220)
    s = _solver(case)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(s.shape)
    y = rng.standard_normal(s.shape)
    scale = np.linalg.norm(x) * np.linalg.norm(y)
    for fwd, bwd in ((s.lddfilterx, s.vjp_filter_x), (s.lddfiltery, s.vjp_filter_y)):
        lhs, rhs = _dot(fwd(x), y), _dot(x, bwd(y))
        print(f"{bwd.__name__}: <Fx,y>={lhs:.15e} <x,F^T y>={rhs:.15e}", flush=True)
        assert abs(lhs - rhs) < 1e-13 * scale
    # ccpml is affine: its linear part is ccpml(x) - ccpml(0)
    Ex = s.ccpml(x) - s.ccpml(np.zeros(s.shape))
    lhs, rhs = _dot(Ex, y), _dot(x, s.vjp_ccpml(y))
    assert abs(lhs - rhs) < 1e-13 * scale
    s.close()


def test_transposed_rhs_against_central_differences():
    case = _euler_case()
    s = _solver(case)
    rng = np.random.default_rng(2)
    U = case.U0
    v = 1e-2 * rng.standard_normal(s.shape)
    lam = rng.standard_normal(s.shape)
    eps = 1e-4
    fd = (s.rhs_euler(U + eps * v) - s.rhs_euler(U - eps * v)) / (2 * eps)
    lhs, rhs = _dot(lam, fd), _dot(s.vjp_rhs(U, lam), v)
    print(f"R'^T: fd {lhs:.12e} adjoint {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) < 1e-8 * abs(lhs)
    s.close()


def test_transposed_step_against_central_differences():
    case = _euler_case(64, 56, seed=72)
    s = _solver(case)
    rng = np.random.default_rng(3)
    U = case.U0
    v = 1e-2 * rng.standard_normal(s.shape)
    lam = rng.standard_normal(s.shape)

    def step(X):
        s.put_state(X)
        s.rk_euler(7)
        return s.get_state()

    eps = 1e-4
    fd = (step(U + eps * v) - step(U - eps * v)) / (2 * eps)
    lhs = _dot(lam, fd)
    rhs = _dot(s.vjp_step(7, U, lam), v)
    print(f"step^T: fd {lhs:.12e} adjoint {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) < 1e-8 * abs(lhs)
    s.close()


def _objective_setup(N=96, D=12, radius=30.0, prec=FP64):
    case = config.adjoint_case(N, D=D, M=0.3, ngauss=4, seed=9, hw_range=(3.0, 6.0), inner=0.5)
    arc = config.observer_arc(N, radius)
    s = _solver(case, prec)
    s.adj_set_arc(arc)
    return case, arc, s


def _J(s, case, dp, nsteps, every=1000):
    U0 = case.Ub.copy()
    U0[..., 3] += dp
    s.put_state(U0)
    return s.adj_forward(0, nsteps, every)


def test_gradient_against_central_differences():
    case, arc, s = _objective_setup()
    dp = case.U0[..., 3] - case.Ub[..., 3]
    nsteps = 24
    J = _J(s, case, dp, nsteps, every=5)
    g = s.adj_backward()[..., 3]
    rng = np.random.default_rng(4)
    for t in range(4):
        d = 1e-3 * rng.standard_normal(dp.shape)
        eps = 1e-3
        fd = (_J(s, case, dp + eps * d, nsteps) - _J(s, case, dp - eps * d, nsteps)) / (2 * eps)
        ad = _dot(g, d)
        print(f"direction {t}: J={J:.6e} fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(fd):.2e}",
              flush=True)
        assert abs(fd - ad) < 1e-6 * abs(fd)
    s.close()


def test_checkpoint_interval_does_not_change_the_gradient():
    case, arc, s = _objective_setup(N=80, D=10, radius=24.0)
    dp = case.U0[..., 3] - case.Ub[..., 3]
    grads = []
    for every in (1, 4, 7, 30):
        J = _J(s, case, dp, 17, every=every)
        grads.append((J, s.adj_backward()))
    for J, g in grads[1:]:
        assert J == grads[0][0]
        assert np.array_equal(g, grads[0][1])
    # the state is the final state again after the backward sweep
    s2 = _solver(case)
    s2.put_state(case.U0)
    s2.run(0, 17)
    assert np.array_equal(s.get_state(), s2.get_state())
    s.close()
    s2.close()


def test_gradient_against_torch_autograd():
    import torch
    from torch_ref import EulerStep
    case, arc, s = _objective_setup(N=41, D=8, radius=12.0)
    dp = case.U0[..., 3] - case.Ub[..., 3]
    nsteps = 3
    J = _J(s, case, dp, nsteps, every=2)
    g = s.adj_backward()[..., 3]
    st = EulerStep(case)
    t = torch.as_tensor(dp).clone().requires_grad_(True)
    Jt, _ = st.objective(t, nsteps, arc)
    (gt,) = torch.autograd.grad(Jt, t)
    gt = gt.numpy()
    err = rel_l2(g, gt)
    print(f"J {J:.15e} torch {float(Jt):.15e}; gradient rel L2 {err:.3e}", flush=True)
    assert abs(J - float(Jt)) < 1e-12 * abs(J)
    assert err < 1e-11
    s.close()


def test_fp32_perturbation_gradient():
    case, arc, s64 = _objective_setup(N=96, D=12, radius=30.0)
    s32 = _solver(case, FP32P)
    s32.adj_set_arc(arc)
    dp = case.U0[..., 3] - case.Ub[..., 3]
    J64 = _J(s64, case, dp, 40, every=10)
    J32 = _J(s32, case, dp, 40, every=10)
    g64 = s64.adj_backward()
    g32 = s32.adj_backward()
    errs = [rel_l2(g32[..., k], g64[..., k]) for k in range(4)]
    print(f"J fp64 {J64:.10e} fp32-pert {J32:.10e}; gradient rel L2 per field {errs}", flush=True)
    assert abs(J32 - J64) < 1e-4 * abs(J64)
    assert errs[3] < 1e-4
    s64.close()
    s32.close()


def test_autograd_function():
    import torch
    from navierstokes_amd.adjoint import observer_objective
    case, arc, s = _objective_setup(N=80, D=10, radius=24.0)
    dp = case.U0[..., 3] - case.Ub[..., 3]
    ref_J = _J(s, case, dp, 12, every=5)
    ref_g = s.adj_backward()[..., 3]
    t = torch.tensor(dp, dtype=torch.float64, device="cuda", requires_grad=True)
    J = observer_objective(t, s, 12, 5)
    J.backward()
    assert float(J) == ref_J
    assert np.array_equal(t.grad.cpu().numpy(), ref_g)
    s.close()


def test_kept_and_recomputed_stage_inputs_give_the_same_gradient():
    """The backward keeps a segment's stage inputs when they fit, else
    recomputes them per step; ns_adj_set_keep(0) forces the second path.
    Both run the same kernels on the same inputs: the same bits."""
    case, arc, s = _objective_setup(N=64, D=8, radius=20.0)
    dp = case.U0[..., 3] - case.Ub[..., 3]
    out = {}
    for keep in (True, False, True):
        s.adj_set_keep(keep)
        J = _J(s, case, dp, 17, every=7)
        out.setdefault(keep, []).append((J, s.adj_backward()))
    (Jk, gk), (Jk2, gk2) = out[True]
    (Jr, gr), = out[False]
    assert Jk == Jr == Jk2
    assert np.array_equal(gk, gr) and np.array_equal(gk, gk2)
    s.close()


def test_two_objectives_on_one_solver():
    """(J1 + J2).backward() with both objectives built on one solver: each
    backward differentiates its own forward (the solver keeps the latest
    forward's checkpoints only, so the earlier objective re-runs its own)."""
    import torch
    from navierstokes_amd.adjoint import observer_objective
    case, arc, s = _objective_setup(N=64, D=8, radius=20.0)
    dp1 = case.U0[..., 3] - case.Ub[..., 3]
    dp2 = 0.5 * np.roll(dp1, 7, axis=1)
    refs = []
    for dp in (dp1, dp2):
        _J(s, case, dp, 10, every=4)
        refs.append(s.adj_backward()[..., 3])
    t1 = torch.tensor(dp1, dtype=torch.float64, device="cuda", requires_grad=True)
    t2 = torch.tensor(dp2, dtype=torch.float64, device="cuda", requires_grad=True)
    J1 = observer_objective(t1, s, 10, 4)
    J2 = observer_objective(t2, s, 10, 4)
    (J1 + J2).backward()
    assert np.array_equal(t1.grad.cpu().numpy(), refs[0])
    assert np.array_equal(t2.grad.cpu().numpy(), refs[1])
    s.close()


def test_adjoint_rejects_viscous_cases():
    from navierstokes_amd._lib import NSError
    case = config.random_case(48, 40, nk=4, D=6, seed=5)
    s = _solver(case)
    with pytest.raises(NSError, match="Euler"):
        s.adj_forward(0, 2, 1)
    s.close()
