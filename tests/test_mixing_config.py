"""Stretched-grid mixing-layer set-up: the map, its metric factor and the
momentum-thickness integral against closed forms."""
import math

import numpy as np

from navierstokes_amd import config


def test_spacing_ratio_and_metric():
    eta, y, dn, meshy = config.tanh_stretched_y(1024, 100.0, 10.0)
    dy = np.diff(y)
    # continuous ratio cosh^2 b = 10; the discrete extreme cells approach it
    # (9.83 at 1024 rows)
    assert abs(dy.max() / dy.min() - 10.0) < 0.25
    assert abs(y[0] + 50.0) < 1e-12 and abs(y[-1] - 50.0) < 1e-12
    # meshy = d eta / dy: compare with a 4th-order difference of the map
    # (its truncation peaks at 1.0e-6 where the map is steepest)
    d4 = (-y[4:] + 8 * y[3:-1] - 8 * y[1:-3] + y[:-4]) / (12 * dn)
    assert np.abs(d4 * meshy[2:-2] - 1.0).max() < 2e-6


def test_momentum_thickness_closed_form():
    # u = (dU/2) tanh(2y/delta): theta = delta/4 for any dU
    yy = np.linspace(-12.0, 12.0, 24001)
    for delta, dU in ((1.0, 0.5), (2.0, 1.3)):
        ub = 0.5 * dU * np.tanh(2.0 * yy / delta)
        th = config.momentum_thickness(ub, yy, 0.5 * dU, -0.5 * dU)
        assert abs(th - delta / 4.0) < 1e-7


def test_mixing_case_base_is_row_constant():
    c = config.mixing_layer_case(128, 96, D=10)
    assert np.array_equal(c.Ub, np.broadcast_to(c.Ub[:, :1, :], c.Ub.shape))
    assert c.metric_everywhere and not c.inviscid
    # Crocco-Busemann temperature at M=0.5 stays within a few percent of 1
    T = c.Ub[..., 3] * 1.4 / c.Ub[..., 0]
    assert np.all(np.abs(T - 1.0) < 0.05)
    # finest cell at the centre equals the x spacing
    y = config.mixing_layer_rows(c)
    assert math.isclose(np.diff(y).min(), c.dm, rel_tol=1e-3)
