"""Batched cases: every case of a batch advances exactly as it would alone.

The batch runs each kernel class as one launch over a work table spanning all
cases (and replays a captured graph per step), but each block computes with
its own case's arguments, so the results must be bitwise those of a
single-case context on the same inputs.  In FP64 every stage variant, of the
batch and of the single context, computes the same bits.  In the FP32
perturbation form (contraction allowed) the tile and marching kernels round
differently (by about 1e-6 of the perturbation after 12 steps), so FP32
comparisons pin both sides to the same kind of kernel: the batch's tile
variant against the single context's tile variant, the batch's marching
variant against the single context's general-region march."""
import numpy as np
import pytest

from conftest import rel_l2
from navierstokes_amd import config

pytestmark = pytest.mark.gpu

FP64, FP32P = 0, 1


def _uneven_pulses(seed):
    return config.batch_sweep_cases(5, seed=seed, sizes=(64, 96, 128, 160), D=12)


def _same(got, ref, case, prec, what):
    assert np.array_equal(got, ref), f"{what}: max diff {np.abs(got - ref).max():.3e}"


def _singles(cases, prec, nsteps, variant=0):
    """Single-case runs: FP64 autotuned (any variant gives the same bits),
    FP32 on the given stage variant (0: tiles, 1: general-region march)."""
    from navierstokes_amd.solver import Solver
    out = []
    for c in cases:
        s = Solver(c, precision=prec)
        if prec == FP32P:
            s.set_variant(variant)
        s.run(0, nsteps)
        out.append(s.get_state())
        s.close()
    return out


@pytest.mark.parametrize("prec", [FP64, FP32P])
def test_batch_equals_single_euler(prec):
    from navierstokes_amd.solver import BatchSolver
    cases = _uneven_pulses(41)
    print(f"[case] batch of {[c.imax for c in cases]} precision={prec}", flush=True)
    ref = _singles(cases, prec, 12)
    b = BatchSolver(cases, precision=prec)
    if prec == FP32P:
        b.set_variant(0)
    b.run(0, 5)
    b.run(5, 7)
    for i, c in enumerate(cases):
        _same(b.get_state(i), ref[i], c, prec, f"case {i} ({c.name})")
    b.close()


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_batch_variants_fp64_bitwise(variant):
    """Every batch stage variant (tiles; marching with fast zones; marching on
    plain general regions) gives the single contexts' FP64 bits, on cases
    from the configured sweep's sizes (128 and 192: look-back chunks, partial
    strips, a fast zone) and uneven small ones (no fast zone, no marching
    region at 64)."""
    from navierstokes_amd.solver import BatchSolver
    cases = _uneven_pulses(45) + config.batch_sweep_cases(3, seed=7, sizes=(128, 192))
    print(f"[case] batch of {[c.imax for c in cases]} FP64 variant {variant}", flush=True)
    ref = _singles(cases, FP64, 7)
    b = BatchSolver(cases, precision=FP64)
    v, ms = b.autotune()
    print(f"batch autotune: variant {v}, ms {ms}", flush=True)
    assert ms[variant] > 0
    b.set_variant(variant)
    b.run(0, 3)
    b.run(3, 4)
    for i, c in enumerate(cases):
        _same(b.get_state(i), ref[i], c, FP64, f"variant {variant} case {i} ({c.name})")
    b.close()


def test_batch_march_fp32_against_single_march():
    """FP32: the batch's marching variant against single contexts on the
    general-region march (variant 1), the same per-point code: equal to
    contraction rounding (reported), p' within 1e-6."""
    from navierstokes_amd.solver import BatchSolver
    cases = _uneven_pulses(46) + config.batch_sweep_cases(2, seed=8, sizes=(192,))
    print(f"[case] batch of {[c.imax for c in cases]} FP32 marching", flush=True)
    ref = _singles(cases, FP32P, 12, variant=1)
    b = BatchSolver(cases, precision=FP32P)
    b.set_variant(1)
    b.run(0, 12)
    for i, c in enumerate(cases):
        got = b.get_state(i)
        err = rel_l2(got[..., 3] - c.Ub[..., 3], ref[i][..., 3] - c.Ub[..., 3])
        print(f"case {i} ({c.name}): bitwise {np.array_equal(got, ref[i])}, p' relative L2 {err:.3e}", flush=True)
        assert err < 1e-6
    b.close()


@pytest.mark.parametrize("nk", [4, 5])
def test_batch_equals_single_viscous(nk):
    from navierstokes_amd.solver import BatchSolver
    cases = [config.random_case(n, m, nk=nk, D=8, seed=50 + n, amp=1e-3, mesh_random=(n % 2 == 0))
             for n, m in ((80, 72), (133, 64), (64, 150))]
    print(f"[case] viscous batch nk={nk}", flush=True)
    for prec in (FP64, FP32P):
        ref = _singles(cases, prec, 6)
        # FP64: every batch variant; FP32: the tile variant, as the singles
        for v in ((0, 1, 2) if prec == FP64 else (0,)):
            b = BatchSolver(cases, precision=prec)
            b.set_variant(v)
            b.run(0, 6)
            for i, c in enumerate(cases):
                _same(b.get_state(i), ref[i], c, prec, f"prec {prec} variant {v} case {i}")
            b.close()


def test_batch_autotune_is_bitwise_neutral():
    """Re-timing the variants (and rebuilding the marching work table) leaves
    the results of the chosen variant unchanged: an FP32 batch that was
    autotuned again gives the bits of a batch pinned to the same variant."""
    from navierstokes_amd.solver import BatchSolver
    cases = _uneven_pulses(43)
    b1 = BatchSolver(cases, precision=FP32P)
    v, ms = b1.autotune()
    print(f"batch autotune: variant {v}, ms {ms}", flush=True)
    assert v in (0, 1, 2) and ms[v] > 0
    b1.run(0, 4)
    b0 = BatchSolver(cases, precision=FP32P)
    b0.set_variant(v)
    b0.run(0, 4)
    for i in range(len(cases)):
        assert np.array_equal(b0.get_state(i), b1.get_state(i))
    b0.close()
    b1.close()


def test_batch_fp32_perturbation_error():
    from navierstokes_amd.solver import BatchSolver
    cases = _uneven_pulses(44)
    b32 = BatchSolver(cases, precision=FP32P)
    b64 = BatchSolver(cases, precision=FP64)
    b32.run(0, 200)
    b64.run(0, 200)
    for i, c in enumerate(cases):
        a, r = b32.get_state(i), b64.get_state(i)
        err = rel_l2(a[..., 3] - c.Ub[..., 3], r[..., 3] - c.Ub[..., 3])
        print(f"case {i} {c.name}: p' relative L2 {err:.3e}", flush=True)
        assert err < 1e-4
    b32.close()
    b64.close()


def test_batch_rejects_mixed_nk():
    from navierstokes_amd._lib import NSError
    from navierstokes_amd.solver import BatchSolver
    cases = [config.acoustic_pulse_case(64, D=8), config.random_case(64, 64, nk=5, D=8, seed=1)]
    with pytest.raises(NSError):
        BatchSolver(cases)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_batch_metric_grids_fp64_bitwise(variant):
    """Batches of cases on non-unit grids (periodic mixing layers: x spacing
    0.05, tanh-stretched y, viscous): the marching variants run the general
    per-point code with 1/h factors on every block (no fast zone); every
    variant gives the single contexts' FP64 bits."""
    from navierstokes_amd.solver import BatchSolver
    cases = [config.mixing_layer_case(160, 96, D=10, seed=5), config.mixing_layer_case(192, 128, D=12, seed=6)]
    print(f"[case] metric-grid batch {[(c.imax, c.jmax) for c in cases]} FP64 variant {variant}", flush=True)
    ref = _singles(cases, FP64, 4)
    b = BatchSolver(cases, precision=FP64)
    b.set_variant(variant)
    b.run(0, 4)
    for i, c in enumerate(cases):
        _same(b.get_state(i), ref[i], c, FP64, f"variant {variant} case {i}")
    b.close()
