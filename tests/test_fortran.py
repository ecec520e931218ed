"""The Fortran side of the boundary (ISO_C_BINDING, amdflang)."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_l2
from navierstokes_amd import config

ROOT = Path(__file__).resolve().parent.parent
FDIR = ROOT / "navierstokes_amd" / "fortran"
DRIVER = FDIR / "ns_run_case"
REF = Path("/root/reference")


def write_case(path, case):
    """Binary case file read by ns_run_case.f90."""
    flags = 1 if case.inviscid else 0
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", case.imax, case.jmax, case.nk, case.D, flags))
        f.write(struct.pack("<8d", case.dm, case.dn, case.dt, case.sigmaf, case.gamma, case.Re, case.Pr,
                            case.Tref))
        for a in (case.meshx, case.meshy, case.sigmax, case.sigmay, case.m, case.n, case.rhob, case.psib,
                  case.Ub, case.U0):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def test_fortran_driver_is_built():
    assert DRIVER.exists(), "build with __graft_entry__.build() (make -C navierstokes_amd/fortran)"


DROPIN_DRIVER = FDIR / "_dropin" / "ns_euler_dropin"


@pytest.mark.skipif(not (REF / "global.f90").exists(), reason="reference tree absent")
def test_dropin_compiles_against_reference_global(tmp_path):
    """ns_dropin.f90 provides the reference's rk/filtering/cc/diff/derivs/
    eqeuler/empirical_coefficients/rhs modules over its unmodified
    global.f90, and the euler.f90-shaped driver links against them."""
    out = subprocess.run(["make", "-s", "-C", str(FDIR), "dropin", f"REF={REF}"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    for mod in ("rk", "filtering", "cc", "diff", "rhs", "derivs", "eqeuler", "empirical_coefficients"):
        assert (FDIR / "_dropin" / f"{mod}.mod").exists(), mod
    assert DROPIN_DRIVER.exists()


def write_dropin_case(path, case, mode, iter0=0, nsteps=0, ct=0.0):
    """Case file read by ns_euler_dropin.f90 (nk = 5: the reference's nVar)."""
    assert case.nk == 5
    with open(path, "wb") as f:
        f.write(struct.pack("<6i", case.imax, case.jmax, case.D, mode, iter0, nsteps))
        f.write(struct.pack("<14d", case.dm, case.dn, case.dt, case.sigmaf, case.gamma, case.Re, case.Pr, case.Tref,
                            case.amplitude, case.wf, case.r0_acous, case.m_pt_acous, case.n_pt_acous, ct))
        for a in (case.meshx, case.meshy, case.sigmax, case.sigmay, case.m, case.n, case.rhob, case.psib,
                  case.Ub, case.U0):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def _source_case():
    case = config.random_case(70, 52, nk=5, D=7, seed=22)
    case.amplitude, case.r0_acous, case.wf = 0.3, 4.0, 0.7
    case.m_pt_acous, case.n_pt_acous = float(case.m[30]), float(case.n[20])
    return case


def _run_dropin(tmp_path, case, mode, **kw):
    fin, fout = tmp_path / f"in{mode}.bin", tmp_path / f"out{mode}.bin"
    write_dropin_case(fin, case, mode, **kw)
    print(f"[case] euler-shaped drop-in driver, mode {mode}, {case.name}", flush=True)
    res = subprocess.run([str(DROPIN_DRIVER), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return np.fromfile(fout, dtype="<f8")


@pytest.mark.gpu
@pytest.mark.skipif(not DROPIN_DRIVER.exists(), reason="drop-in driver not built (needs the reference's global.f90)")
def test_euler_shaped_driver_through_dropin_modules(tmp_path):
    """The reference's call contract through the drop-in modules, run by a
    driver shaped like euler.f90 over the reference's own global.f90:
    rk_euler in the time loop (euler.f90:123-130) for 3 steps with the
    acoustic source on; acousticfont then rhs_euler (rk.f90:55-58: the RHS
    sees the source acousticfont just set); rhs_euler's body through
    MOLECULAR_MI, k_coefficient, deriv, tau_ij and EULERNONLINEAR; metricsX
    and metricsY.  Each against the oracle (or the metric's closed form)."""
    from oracle.oracle import Oracle
    case = _source_case()
    o = Oracle(case)
    shape = (case.jmax, case.imax, case.nk)
    U = _run_dropin(tmp_path, case, 1, iter0=0, nsteps=3).reshape(shape)
    Uo = o.run(case.U0, 0, 3)
    err = rel_l2(U, Uo)
    print(f"3 steps (rk_euler): rel L2 vs oracle {err:.3e}", flush=True)
    assert err < 1e-13
    ref = o.rhs(case.U0, o.acousticfont(0.4, 17))
    assert np.abs(o.acousticfont(0.4, 17)).max() > 1e-3
    for mode in (2, 3):
        d = _run_dropin(tmp_path, case, mode, iter0=17, ct=0.4).reshape(shape)
        e = max(np.abs(d[..., k] - ref[..., k]).max() / np.abs(ref[..., k]).max() for k in range(case.nk))
        print(f"mode {mode}: RHS max rel err vs oracle {e:.3e}", flush=True)
        assert e < 1e-12
    m = _run_dropin(tmp_path, case, 4).reshape((2,) + shape)
    i1, j1 = np.arange(1, case.imax + 1), np.arange(1, case.jmax + 1)
    metx = np.where((i1 <= case.D) | (i1 >= case.imax - case.D + 1), case.meshx, 1.0)
    mety = np.where((j1 <= case.D) | (j1 >= case.jmax - case.D + 1), case.meshy, 1.0)
    assert np.array_equal(m[0], case.U0 * metx[None, :, None])
    assert np.array_equal(m[1], case.U0 * mety[:, None, None])


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    from oracle.oracle import Oracle
    case = config.random_case(70, 52, nk=5, D=7, seed=21)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_case(fin, case)
    print(f"[case] fortran driver {case.name} 3 steps", flush=True)
    res = subprocess.run([str(DRIVER), str(fin), str(fout), "3"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    Ug = np.fromfile(fout, dtype="<f8").reshape(case.jmax, case.imax, case.nk)
    Uo = Oracle(case).run(case.U0, 0, 3)
    assert rel_l2(Ug, Uo) < 1e-13
