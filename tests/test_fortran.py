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


@pytest.mark.skipif(not (REF / "global.f90").exists(), reason="reference tree absent")
def test_dropin_compiles_against_reference_global(tmp_path):
    """ns_dropin.f90 provides the reference's rk/filtering/cc/diff/rhs modules
    over its unmodified global.f90."""
    out = subprocess.run(["make", "-s", "-C", str(FDIR), "dropin", f"REF={REF}"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    for mod in ("rk", "filtering", "cc", "diff", "rhs"):
        assert (FDIR / "_dropin" / f"{mod}.mod").exists(), mod


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
