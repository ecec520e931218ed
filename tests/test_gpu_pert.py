"""FP32 perturbation form against the FP64 path on the same inputs.

The bar is the one the perturbation form is specified with: relative L2 of
the pressure perturbation p' = p - p_b within 1e-4 of FP64."""
import numpy as np
import pytest

from conftest import rel_l2
from navierstokes_amd import config

pytestmark = pytest.mark.gpu

FP64, FP32P = 0, 1


def _solver(case, prec):
    from navierstokes_amd.solver import Solver
    print(f"[case] {case.name} {case.imax}x{case.jmax} nk={case.nk} precision={prec}", flush=True)
    return Solver(case, precision=prec)


def _pert_err(a, b, base, k):
    return rel_l2(a[..., k] - base[..., k], b[..., k] - base[..., k])


def test_state_roundtrip_and_probe():
    case = config.random_case(64, 48, nk=5, D=6, seed=31)
    s = _solver(case, FP32P)
    U = s.get_state()
    assert np.abs(U - case.U0).max() < 1e-7 * np.abs(case.U0 - case.Ub).max() + 1e-15
    v = s.probe([(3, 4), (64, 48)])
    # q = U - Ub is stored in FP32: half an ulp of |q| <= 1
    assert np.allclose(v[0], case.U0[3, 2], rtol=0, atol=1e-7)
    s.check_finite()
    s.close()


def test_rejects_base_flow_varying_along_rows():
    from navierstokes_amd._lib import NSError
    case = config.random_case(48, 40, nk=4, D=6, seed=32)
    case.Ub = case.Ub.copy()
    case.Ub[5, 7, 0] += 1e-3
    with pytest.raises(NSError, match="constant along each row"):
        _solver(case, FP32P)


def test_base_state_stays_at_rest():
    case = config.acoustic_pulse_case(201)
    case.U0 = case.Ub.copy()
    s32, s64 = _solver(case, FP32P), _solver(case, FP64)
    s32.run(0, 20)
    s64.run(0, 20)
    a, b = s32.get_state(), s64.get_state()
    # only the float32 y-literal drift moves the base state; both forms agree on it
    assert np.abs(a - b).max() < 1e-12
    s32.close()
    s64.close()


@pytest.mark.parametrize("nsteps", [100, 1000])
def test_pulse_201(nsteps):
    case = config.acoustic_pulse_case(201)
    s32, s64 = _solver(case, FP32P), _solver(case, FP64)
    s32.run(0, nsteps)
    s64.run(0, nsteps)
    a, b = s32.get_state(), s64.get_state()
    err = _pert_err(a, b, case.Ub, 3)
    print(f"p' relative L2 after {nsteps} steps: {err:.3e}", flush=True)
    assert err < 1e-4
    if nsteps <= 200:
        # while the pulse is inside the grid every perturbation field meets the bar;
        # by step 1000 it has left and u, v are a 1e-8 residual
        for k in range(3):
            assert _pert_err(a, b, case.Ub, k) < 1e-4
    s32.close()
    s64.close()


@pytest.mark.parametrize("nk", [4, 5])
def test_viscous_mixing_layer_base(nk):
    case = config.random_case(160, 120, nk=nk, D=12, seed=33, amp=1e-3, mesh_random=False)
    s32, s64 = _solver(case, FP32P), _solver(case, FP64)
    s32.run(0, 20)
    s64.run(0, 20)
    a, b = s32.get_state(), s64.get_state()
    for k in range(nk):
        err = _pert_err(a, b, case.Ub, k)
        print(f"k={k} perturbation relative L2: {err:.3e}", flush=True)
        assert err < 1e-4
    s32.close()
    s64.close()


def test_pulse_vortex_ns_512():
    case = config.pulse_vortex_case(512, D=40)
    s32, s64 = _solver(case, FP32P), _solver(case, FP64)
    s32.run(0, 50)
    s64.run(0, 50)
    a, b = s32.get_state(), s64.get_state()
    err = _pert_err(a, b, case.Ub, 3)
    print(f"p' relative L2 (NS pulse+vortex 512^2, 50 steps): {err:.3e}", flush=True)
    assert err < 1e-4
    s32.close()
    s64.close()


@pytest.mark.parametrize("nk", [4, 5])
def test_stage_variants_one_context(nk):
    """Every stage variant of the perturbation form (interior tiles; marching
    kernel over the general region; marching kernel over the fast zone),
    switched in one context from the same initial state, stays within the
    perturbation bar of FP64, and the variants agree with one another to
    FP32 rounding."""
    case = config.random_case(330, 212, nk=nk, D=20, seed=34, amp=1e-3)
    s64 = _solver(case, FP64)
    s64.run(0, 10)
    b = s64.get_state()
    s64.close()
    s = _solver(case, FP32P)
    ref = None
    for v in (0, 1, 2):
        print(f"[variant] {v}", flush=True)
        s.put_state(case.U0)
        s.set_variant(v)
        s.run(0, 10)
        a = s.get_state()
        errs = [_pert_err(a, b, case.Ub, k) for k in range(nk)]
        print(f"variant {v}: perturbation relative L2 per field {errs}", flush=True)
        assert max(errs) < 1e-4
        if ref is None:
            ref = a
        else:
            assert np.abs(a - ref).max() <= 1e-5 * np.abs(ref - case.Ub).max()
    s.close()


@pytest.mark.parametrize("nk", [4, 5])
def test_stage_variants(nk):
    """Every stage variant of the perturbation form (interior tiles; marching
    kernel over the general region; marching kernel over the fast zone)
    stays within the perturbation bar of FP64, and the variants agree with
    one another to FP32 rounding."""
    case = config.random_case(330, 212, nk=nk, D=20, seed=34, amp=1e-3)
    s64 = _solver(case, FP64)
    s64.run(0, 10)
    b = s64.get_state()
    s64.close()
    ref = None
    for v in (0, 1, 2):
        s = _solver(case, FP32P)
        s.set_variant(v)
        s.run(0, 10)
        a = s.get_state()
        s.close()
        errs = [_pert_err(a, b, case.Ub, k) for k in range(nk)]
        print(f"variant {v}: perturbation relative L2 per field {errs}", flush=True)
        assert max(errs) < 1e-4
        if ref is None:
            ref = a
        else:
            assert np.abs(a - ref).max() <= 1e-5 * np.abs(ref - case.Ub).max()
