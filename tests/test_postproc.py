"""Growth-rate post-processing (fftdns.py:42-165) against closed forms: for
station energies E_l(t) = E0 exp(-2 alpha x_l) g(t), the zero-frequency
component is proportional to exp(-2 alpha x_l), so all three formulas return
alpha exactly; and H0 is the plain sum of the series (numpy's FFT)."""
import math

import numpy as np

from navierstokes_amd.postproc import growth_rates


def _series(alpha, dx, nt=400, ns=10, dt=0.01):
    t = np.arange(nt) * dt
    x = 10.0 + np.arange(ns) * dx
    g = 1.0 + 0.3 * np.sin(2.0 * math.pi * 3.0 * t) + 0.1 * np.cos(5.0 * t)
    return np.exp(-2.0 * alpha * x)[None, :] * g[:, None]


def test_growth_rates_closed_form():
    for alpha in (0.05, -0.12, 0.3):
        E = _series(alpha, 0.05)
        r = growth_rates(E, dx=0.05, dt=0.01, l=5, freq=2.0)
        for k in ("alpha1", "alpha3", "alpha4"):
            assert abs(r[k] - alpha) < 1e-10 * max(1.0, abs(alpha)), (k, r[k], alpha)
        assert np.allclose(r["H0"], E.sum(axis=0), rtol=1e-13)
        assert abs(r["bin"] * r["dw"] - 2.0) <= r["dw"] / 2 + 1e-12


def test_growth_rates_reject_short_station_lists():
    import pytest
    with pytest.raises(ValueError):
        growth_rates(_series(0.1, 0.05, ns=6), dx=0.05, dt=0.01, l=5)
