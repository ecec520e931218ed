"""The float64 torch restatement used as the autograd reference for the
adjoint kernels: its forward step must be the oracle's (which is pinned to
the reference), so its autograd gradient is the gradient of the reference's
discrete time advance."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from navierstokes_amd import config
from torch_ref import EulerStep, aos


def _cases():
    a = config.acoustic_pulse_case(41, D=8, u0=0.3, half_width=3.0)
    b = config.random_case(44, 38, nk=4, D=7, seed=61, amp=1e-3, inviscid=True, mesh_random=True, Re=1e300)
    return [a, b]


@pytest.mark.parametrize("which", [0, 1])
def test_torch_step_matches_oracle(which):
    from oracle.oracle import Oracle
    case = _cases()[which]
    ref = Oracle(case).run(case.U0.copy(), 0, 3)
    st = EulerStep(case)
    U = torch.as_tensor(case.U0).permute(2, 0, 1).contiguous()
    for _ in range(3):
        U = st.step(U)
    got = aos(U)
    for k in range(4):
        d = case.U0[..., k] - case.Ub[..., k]
        err = rel_l2(got[..., k] - case.Ub[..., k], ref[..., k] - case.Ub[..., k]) if np.abs(d).max() > 0 else \
            np.abs(got[..., k] - ref[..., k]).max()
        assert err < 1e-11, (k, err)


def test_torch_gradient_matches_finite_differences():
    case = config.acoustic_pulse_case(33, D=6, u0=0.2)
    st = EulerStep(case)
    arc = config.observer_arc(33, 10.0)
    rng = np.random.default_rng(3)
    dp0 = torch.as_tensor(1e-3 * rng.standard_normal((33, 33)))
    dp = dp0.clone().requires_grad_(True)
    J, _ = st.objective(dp, 2, arc)
    (g,) = torch.autograd.grad(J, dp)
    d = torch.as_tensor(rng.standard_normal((33, 33)))
    # central differences converge as eps^2 (5e-6 relative at 1e-5, 6e-9 at
    # 1e-6); below that the rounding of the states in J takes over
    eps = 1e-6
    Jp, _ = st.objective(dp0 + eps * d, 2, arc)
    Jm, _ = st.objective(dp0 - eps * d, 2, arc)
    fd = float((Jp - Jm) / (2 * eps))
    assert abs(fd - float((g * d).sum())) < 1e-7 * abs(fd)
