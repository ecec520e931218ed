"""Periodic x (NS_FLAG_PERIODIC_X; no reference counterpart) on the GPU
against the oracle's periodic mode: wrapped x derivatives, the ring x filter
(look-back chunks; the chunk holding the end recomputes the sweep's head),
ccpml rows on every column and no inflow/outflow columns, y buffers only,
the LDDRK update on every column, and whole steps in each stage variant
(the marching kernel covers every column with wrapped windows)."""
import numpy as np
import pytest

from conftest import rel_l2, rel_max
from navierstokes_amd import config
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64, FP32P = 0, 1

CASES = [
    dict(imax=96, jmax=40, nk=4, D=6, seed=31),
    dict(imax=700, jmax=300, nk=5, D=12, seed=32),    # several ring-filter chunks
    dict(imax=333, jmax=157, nk=4, D=9, seed=33),     # partial last marching strip
]


def _case(i, **kw):
    c = config.random_case(**{**CASES[i], **kw})
    c.periodic_x = True
    return c


def _solver(case, prec=FP64):
    from navierstokes_amd.solver import Solver
    print(f"[case] periodic {case.name} {case.imax}x{case.jmax} nk={case.nk} precision={prec}", flush=True)
    return Solver(case, precision=prec)


@pytest.mark.parametrize("i", range(len(CASES)))
def test_periodic_operators(i):
    case = _case(i)
    s, o = _solver(case), Oracle(case)
    U = case.U0
    assert rel_max(s.deropm(U), o.deropm(U)) < 1e-14
    fx, ox = s.lddfilterx(U), o.filterx(U)
    assert rel_max(fx - U, ox - U) < 1e-12 and rel_max(fx, ox) < 1e-15
    assert np.array_equal(s.ccpml(U), o.ccpml(U))
    assert np.array_equal(s.bufferxy(U), o.bufferxy(U))
    rng = np.random.default_rng(9)
    dudt, w = rng.standard_normal(U.shape), rng.standard_normal(U.shape)
    Ug, wg = s.rk_steep(-0.7, 0.8, U, dudt, w)
    Uo, wo = o.rk_steep(-0.7, 0.8, U, dudt, w)
    assert rel_max(Ug, Uo) < 1e-15 and rel_max(wg, wo) < 1e-15
    for inviscid in (False, True):
        case.inviscid, case.Re = inviscid, (1e300 if inviscid else 500.0)
        s2, o2 = _solver(case), Oracle(case)
        dg, do = s2.rhs_euler(U), o2.rhs(U)
        for k in range(case.nk):
            assert rel_max(dg[..., k], do[..., k]) < 1e-12, (inviscid, k)
        s2.close()
    s.close()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_periodic_steps_every_variant(i):
    case = _case(i)
    nsteps = 2
    Uo = Oracle(case).run(case.U0, 0, nsteps)
    ref = None
    for v in (0, 1, 2):
        s = _solver(case)
        s.set_variant(v)
        s.run(0, nsteps)
        U = s.get_state()
        s.close()
        errs = [rel_l2(U[..., k], Uo[..., k]) for k in range(case.nk)]
        print(f"variant {v}: rel L2 vs oracle {max(errs):.3e}", flush=True)
        assert max(errs) < 1e-13
        if ref is None:
            ref = U
        else:
            assert np.array_equal(U, ref), f"variant {v}: max diff {np.abs(U - ref).max():.3e}"


def test_periodic_unit_grid_fast_code():
    """A unit grid: the fast per-point code runs in every column (no x
    layers), the general code in the y layers."""
    case = _case(1, mesh_random=False)
    case.dm = case.dn = 1.0
    Uo = Oracle(case).run(case.U0, 0, 3)
    for v in (0, 2):
        s = _solver(case)
        s.set_variant(v)
        s.run(0, 3)
        U = s.get_state()
        s.close()
        err = max(rel_l2(U[..., k], Uo[..., k]) for k in range(case.nk))
        print(f"variant {v}: rel L2 vs oracle {err:.3e}", flush=True)
        assert err < 1e-13


def test_periodic_fp32_perturbation():
    case = config.mixing_layer_case(256, 160, D=12)
    assert case.periodic_x
    s64, s32 = _solver(case, FP64), _solver(case, FP32P)
    s64.run(0, 200)
    s32.run(0, 200)
    a, b = s32.get_state(), s64.get_state()
    err = rel_l2(a[..., 2], b[..., 2])
    print(f"mixing layer 256x160 periodic, 200 steps: v relative L2 FP32-pert vs FP64 {err:.3e}", flush=True)
    assert err < 1e-4
    s64.close()
    s32.close()


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_periodic_batch_equals_single(variant):
    """Every batch stage variant (tiles; marching, whose strips wrap around
    the period) gives the single contexts' FP64 bits."""
    from navierstokes_amd.solver import BatchSolver, Solver
    cases = [_case(0), _case(2)]   # both nk = 4
    print(f"[case] periodic batch, variant {variant}", flush=True)
    b = BatchSolver(cases)
    b.set_variant(variant)
    b.run(0, 3)
    for i, c in enumerate(cases):
        s = Solver(c)
        s.run(0, 3)
        assert np.array_equal(b.get_state(i), s.get_state())
        s.close()
    b.close()
