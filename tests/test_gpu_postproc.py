"""Post-processing of a forced run on station data from the GPU.

The reference's own forced configuration (dados.in defaults: the mixing
layer with the acoustic source at amplitude 0.8, wf = 1, init.f90 and
rk.f90:171-193) on a reduced 201 x 161 grid: ten consecutive station
columns sampled on the device during the run (ns_station_energy,
outt.f90:63-110), growth rates formed on the GPU by the torch restatement
of fftdns.py:42-165 (postproc.growth_rates).  The same pipeline on the
oracle's states (its stations summed in numpy as outt.f90:92-110 writes
them) gives the same energies and rates.
"""
import numpy as np
import pytest
import torch

from navierstokes_amd import config
from navierstokes_amd.postproc import growth_rates, record_stations

pytestmark = pytest.mark.gpu

NSTEPS, EVERY = 400, 4
COLS = list(range(80, 90))  # 1-based station columns


def _stations(U, case, cols):
    D, jm, dn = case.D, case.jmax, case.dn
    out = []
    for i in cols:
        u, v = U[:, i - 1, 1], U[:, i - 1, 2]
        k = 0.5 * (u * u + v * v)
        out.append((2.0 * k[D:jm - D - 1].sum() + (k[D - 1] + k[jm - D - 1])) * dn / 2.0)
    return out


def _case():
    return config.reference_case({"imax": 201, "jmax": 161, "d": 30})


def test_forced_run_growth_rates_against_the_oracle():
    from navierstokes_amd.solver import Solver
    from oracle.oracle import Oracle

    case = _case()
    print(f"[case] {case.name} {case.imax}x{case.jmax} nk={case.nk} amplitude={case.amplitude} wf={case.wf}",
          flush=True)
    with Solver(case) as s:
        t, E = record_stations(s, COLS, 0, NSTEPS, EVERY)
    rates = growth_rates(torch.as_tensor(E, device="cuda"), dx=case.dm, dt=case.dt * EVERY, l=5, freq=case.wf)

    o = Oracle(case, parallel=True)
    U = case.U0.copy()
    Eo = [_stations(U, case, COLS)]
    for n in range(0, NSTEPS, EVERY):
        U = o.run(U, n, EVERY)
        Eo.append(_stations(U, case, COLS))
    Eo = np.asarray(Eo)
    ro = growth_rates(Eo, dx=case.dm, dt=case.dt * EVERY, l=5, freq=case.wf)

    assert E.shape == Eo.shape == (NSTEPS // EVERY + 1, len(COLS))
    assert np.allclose(t, np.arange(NSTEPS // EVERY + 1) * EVERY * case.dt)
    e_err = np.abs(E - Eo).max() / np.abs(Eo).max()
    change = np.abs(Eo[-1] - Eo[0]).max() / np.abs(Eo[0]).max()
    print(f"station energies: max relative difference {e_err:.3e}; change over the run {change:.3e}", flush=True)
    assert change > 1e-6  # the source has driven the field
    assert e_err < 1e-11
    for key in ("alpha1", "alpha3", "alpha4"):
        print(f"{key}: GPU {rates[key]:.12e} oracle {ro[key]:.12e}", flush=True)
        assert np.isfinite(rates[key])
        assert abs(rates[key] - ro[key]) <= 1e-8 * max(abs(ro[key]), 1e-12)
    assert rates["bin"] == ro["bin"]


def test_forced_run_growth_rates_fp32_perturbation():
    """The same stations from the FP32 perturbation form: its growth rates
    stay within 1e-4 of the FP64 run's."""
    from navierstokes_amd.solver import Solver

    case = _case()
    out = {}
    for prec in (0, 1):
        with Solver(case, precision=prec) as s:
            _, E = record_stations(s, COLS, 0, NSTEPS, EVERY)
        out[prec] = growth_rates(E, dx=case.dm, dt=case.dt * EVERY, l=5, freq=case.wf, device="cuda")
    for key in ("alpha1", "alpha3", "alpha4"):
        err = abs(out[1][key] - out[0][key]) / abs(out[0][key])
        print(f"{key}: FP64 {out[0][key]:.6e} FP32-pert {out[1][key]:.6e} rel diff {err:.2e}", flush=True)
        assert err < 1e-4
