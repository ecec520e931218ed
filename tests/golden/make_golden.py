"""Write golden vectors from the reference's own compiled routines.

oracle/_ref/libnsref.so is built (oracle/Makefile, target `ref`) from the
reference tree's unmodified global.f90, coefficients.f90, cc.f90 and
diffall.f90.  This script feeds it seeded inputs and stores inputs and
outputs in ref_vectors.npz, so the oracle stays pinned where the reference
tree is absent.  Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import Reference  # noqa: E402

OUT = Path(__file__).resolve().parent / "ref_vectors.npz"


def main():
    rng = np.random.default_rng(20240611)
    imax, jmax, dm, dn, gamma, Tref = 23, 19, 0.37, 0.53, 1.4, 273.15
    ref = Reference(imax, jmax, dm, dn, gamma, Tref)
    # smooth + noise field, nk = 5, layout (jmax, imax, nk) = U(1:5,1:imax,1:jmax)
    X, Y = np.meshgrid(np.arange(imax) * dm, np.arange(jmax) * dn)
    U = np.stack([1.0 + 0.1 * np.sin(0.7 * X + 0.3 * k) * np.cos(0.4 * Y - 0.2 * k)
                  + 0.01 * rng.standard_normal((jmax, imax)) for k in range(5)], axis=-1)
    dUdm = ref.deropm(U)
    dUdn = ref.deropn(U)
    T = 0.6 + 0.8 * rng.random((jmax, imax))
    mi, kk = ref.transport(T)
    coef = ref.get_coefficients()
    rhob = 1.0 + 0.1 * rng.random(jmax)
    psib = rng.random(jmax)
    Ucc = ref.ccpml(U, rhob, psib)
    np.savez_compressed(
        OUT, imax=imax, jmax=jmax, dm=dm, dn=dn, gamma=gamma, Tref=Tref, U=U, dUdm=dUdm, dUdn=dUdn,
        T=T, mi=mi, k=kk, coef_names=np.array(list(coef.keys())),
        coef_values=np.array(list(coef.values())), rhob=rhob, psib=psib, U_ccpml=Ucc)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
