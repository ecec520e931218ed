"""Snapshot writers (host side): column order and layout of the text file
against the reference's output loop (outt.f90:49-55), with a stand-in for
the solver's device calls."""
import numpy as np

from navierstokes_amd import config
from navierstokes_amd.output import write_contorno, write_snapshot


class _Stub:
    def __init__(self, case):
        self.case = case
        rng = np.random.default_rng(0)
        self.U = rng.standard_normal((case.jmax, case.imax, case.nk))
        self.om = rng.standard_normal((case.jmax, case.imax))

    def get_state(self):
        return self.U

    def vorticity(self):
        return self.om


def test_contorno_layout(tmp_path):
    case = config.random_case(20, 17, nk=5, D=4, seed=2)
    st = _Stub(case)
    p = tmp_path / "contorno_1.dat"
    write_contorno(p, st)
    blocks = p.read_text().strip("\n").split("\n\n")
    assert len(blocks) == case.imax
    i, j = 7, 11
    row = np.array(blocks[i].split("\n")[j].split(), dtype=float)
    expect = [case.m[i], case.n[j], *st.U[j, i, :4], st.om[j, i], st.U[j, i, 4]]
    assert np.array_equal(row, np.array(expect))


def test_npz_snapshot(tmp_path):
    case = config.random_case(24, 18, nk=4, D=4, seed=3)
    st = _Stub(case)
    p = tmp_path / "snap.npz"
    write_snapshot(p, st)
    z = np.load(p)
    assert np.array_equal(z["U"], st.U) and np.array_equal(z["omega"], st.om)
    assert np.array_equal(z["m"], case.m) and np.array_equal(z["n"], case.n)
