"""Snapshots of the state: the output path of the reference's time loop.

outt.f90:16-57 (`output`) writes, every few hundred steps, an ASCII file
contorno_N.dat with one line per grid point -- m(i), n(j), U(1..4,i,j), the
vorticity dUcdm(3) - dUcdn(2) and U(5,i,j) -- i outer, j inner, a blank line
after each i (gnuplot blocks).  Here the vorticity comes from the device
(`Solver.vorticity`, the diffall.f90 operators) and the state from
`Solver.get_state`; `write_snapshot` stores the same content in binary
(numpy .npz), `write_contorno` as text in the reference's column order.
"""
from __future__ import annotations

import numpy as np


def snapshot(solver) -> dict:
    c = solver.case
    return {"m": np.asarray(c.m, dtype=np.float64), "n": np.asarray(c.n, dtype=np.float64),
            "U": solver.get_state(), "omega": solver.vorticity()}


def write_snapshot(path, solver) -> dict:
    """Binary snapshot: m (imax), n (jmax), U (jmax, imax, nk), omega (jmax, imax)."""
    snap = snapshot(solver)
    np.savez(path, **snap)
    return snap


def write_contorno(path, solver, snap: dict | None = None) -> None:
    """Text snapshot in the column order of outt.f90:51-55 (free format)."""
    s = snap or snapshot(solver)
    m, n, U, om = s["m"], s["n"], s["U"], s["omega"]
    jmax, imax, nk = U.shape
    with open(path, "w") as f:
        for i in range(imax):
            cols = [np.full(jmax, m[i]), n, U[:, i, 0], U[:, i, 1], U[:, i, 2], U[:, i, 3], om[:, i]]
            if nk >= 5:
                cols.append(U[:, i, 4])
            np.savetxt(f, np.column_stack(cols), fmt="%.16e")
            f.write("\n")
