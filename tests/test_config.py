"""Host-side set-up (init.f90 restatement and workload builders)."""
import math
from pathlib import Path

import numpy as np
import pytest

from navierstokes_amd import config

REF_DADOS = Path("/root/reference/dados.in")


def test_grid_profiles_closed_forms():
    g = config.grid_profiles(201, 161, -100.0, 100.0, -80.0, 80.0, 30, 0.01, 20.0)
    assert g["dm"] == 1.0 and g["dn"] == 1.0
    # stretch is 1 at the buffer interface and 1/(1+2) at the outer edge
    assert g["meshx"][30] == 1.0 and g["meshy"][30] == 1.0
    assert abs(g["meshx"][0] - 1.0 / 3.0) < 1e-15 and abs(g["meshx"][-1] - 1.0 / 3.0) < 1e-15
    assert np.all(g["meshx"][31:-31] == 1.0)
    # Wasistho profile: 0 at the outer edge, -> 1 towards the interior
    assert g["sigmax"][-1] == 0.0
    assert 0.98 < g["sigmax"][-30] <= 1.0
    assert np.all(np.diff(g["sigmax"][-30:]) <= 0)
    # the reference's imaxpml/jmaxpml are never set: left/bottom layers stay 0
    assert np.all(g["sigmax"][:30] == 0.0) and np.all(g["sigmay"][:30] == 0.0)
    g2 = config.grid_profiles(201, 161, -100.0, 100.0, -80.0, 80.0, 30, 0.01, 20.0, sponge_left_bottom=True)
    assert np.allclose(g2["sigmax"][:30], g2["sigmax"][::-1][:30])


def test_mixing_layer_base_flow():
    n = np.linspace(-2.4, 2.4, 101)
    rhob, psib, Ub = config.mixing_layer_fields(n, 7, 5, 0.9, 0.1, 1.0, 0.5)
    assert abs(Ub[-1, 0, 1] - 0.9) < 1e-6 and abs(Ub[0, 0, 1] - 0.1) < 1e-6
    assert np.allclose(Ub[..., 3], 1 / 1.4)
    # Crocco-Busemann with equal pressure: rho = 1/T
    assert abs(rhob[-1] - 1.0) < 1e-6 and abs(rhob[0] - 2.0) < 1e-5
    assert psib[0] < 1e-6 and psib[-1] > 1 - 1e-6


@pytest.mark.skipif(not REF_DADOS.exists(), reason="reference tree absent")
def test_read_reference_dados():
    d = config.read_dados(REF_DADOS)
    assert d["imax"] == 521 and d["jmax"] == 481 and d["d"] == 90
    assert d["dt"] == 0.01 and d["sigmaf"] == 0.1 and d["tref"] == 273.15
    case = config.reference_case(d)
    assert case.imax == 521 and abs(case.dm - 26.0 / 520) < 1e-15


def test_acoustic_pulse_case():
    c = config.acoustic_pulse_case(201)
    assert c.dm == 1.0 and c.dt == 0.5 and c.D == 30 and c.sigmaf == 0.2
    p0 = 1 / 1.4
    assert abs(c.U0[100, 100, 3] - p0 * (1 + 1e-3)) < 1e-15
    # isentropic density and half-width 3
    assert abs(c.U0[100, 100, 0] - (1 + 1e-3) ** (1 / 1.4)) < 1e-15
    assert abs(c.U0[100, 103, 3] - p0 - 0.5e-3 * p0) < 1e-15
    assert np.all(c.U0[..., 1] == 0)


def test_pulse_vortex_case_equilibrium():
    c = config.pulse_vortex_case(513, D=40)   # odd: the row j=256 is y=0
    assert abs(c.dt - 1 / 3) < 1e-15 and abs(c.Re - 2500.0) < 1e-9
    # radial equilibrium of the isentropic vortex: dp/dr = rho u_theta^2 / r
    j = 256
    xs = c.m
    i0 = int(np.argmin(np.abs(xs + 200.0)))
    r = xs[i0 + 1:i0 + 25] - xs[i0]
    p = c.U0[j, i0 + 1:i0 + 25, 3]
    rho = c.U0[j, i0 + 1:i0 + 25, 0]
    ut = c.U0[j, i0 + 1:i0 + 25, 2]
    dpdr = np.gradient(p, r)
    # (np.gradient is 2nd order at spacing R/10: residual ~2%)
    assert np.allclose(dpdr[2:-2], (rho * ut ** 2 / r)[2:-2], rtol=3e-2)
