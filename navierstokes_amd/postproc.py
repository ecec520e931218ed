"""Spatial growth rates from station energies, in torch (the post-processing
of the reference's forced runs, fftdns.py:42-165).

During a run the reference writes, at each output time, the kinetic energy
of ten consecutive columns integrated over y (outt.f90:63-110); here
`Solver.station_energy` computes them on the device and `record_stations`
collects the time series.  fftdns.py then takes the FFT in time of each
station's series, keeps the real part, and forms growth rates from the
zero-frequency component H0 between stations l and its neighbours:
    alpha1 = -1/2 ln(H0[l+1] / H0[l]) / dx                  (forward)
    alpha3 = -1/2 ln(H0[l+1] / H0[l-1]) / (2 dx)            (central)
    alpha4 = -1/2 ln(H0[l+3]^8 H0[l] / (H0[l+4] H0[l+1]^8)) / (12 dx)
with l the 1-based station index (l = 5 in the reference), and reports the
frequency bin nearest the forcing frequency (dw = pi / (dt N) as written
there).
"""
from __future__ import annotations

import math

import numpy as np
import torch


def record_stations(solver, cols, iter0: int, nsteps: int, every: int) -> tuple[np.ndarray, np.ndarray]:
    """Advance `nsteps` from `iter0`, sampling the station energies every
    `every` steps (and at the start).  Returns (time, E) with E of shape
    (nsamples, nstations)."""
    t, E = [], []
    n = 0
    while True:
        E.append(solver.station_energy(cols))
        t.append((iter0 + n) * solver.case.dt)
        if n >= nsteps:
            break
        k = min(every, nsteps - n)
        solver.run(iter0 + n, k)
        n += k
    return np.asarray(t), np.asarray(E)


def growth_rates(E, dx: float, dt: float, l: int = 5, freq: float | None = None, device=None) -> dict:
    """Growth rates of fftdns.py from the station series E (nsamples x
    nstations, stations 1-based in the formulas).  Works on the device given
    (or E's own when E is a tensor)."""
    H = torch.as_tensor(E, dtype=torch.float64, device=device)
    nt, ns = H.shape
    if not 2 <= l <= ns - 4:
        raise ValueError(f"l = {l} needs stations l-1 .. l+4 (have {ns})")
    Hf = torch.fft.fft(H, dim=0).real            # per station, as fftdns.py:70-74
    H0 = Hf[0]                                   # zero-frequency component
    h = lambda s: H0[s - 1]                      # 1-based station index
    a1 = -0.5 * torch.log(h(l + 1) / h(l)) / dx
    a3 = -0.5 * torch.log(h(l + 1) / h(l - 1)) / (2.0 * dx)
    a4 = -0.5 * torch.log(h(l + 3) ** 8 * h(l) / (h(l + 4) * h(l + 1) ** 8)) / (12.0 * dx)
    out = {"alpha1": float(a1), "alpha3": float(a3), "alpha4": float(a4), "H0": H0.cpu().numpy()}
    if freq is not None:
        dw = math.pi / (dt * nt)
        kf = torch.arange(nt, dtype=torch.float64, device=H.device) * dw
        out["bin"] = int(torch.argmin(torch.abs(kf - freq)))
        out["dw"] = dw
    return out
