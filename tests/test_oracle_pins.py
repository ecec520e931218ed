"""Pins of the CPU oracle (oracle/ns_oracle.c) that do not rest on the oracle
itself: the reference's own compiled routines (bitwise), golden vectors
written from them, closed forms and independent formulations."""
from pathlib import Path

import numpy as np
import pytest
from scipy.linalg import solve_triangular

from conftest import rel_l2, rel_max
from navierstokes_amd import config
from oracle.oracle import Oracle, Reference

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_vectors.npz"

# filtering.f90:54-66 and :25-31
DF = np.array([0.001254597714, -0.008520738659, 0.029662754736, -0.069975429105, 0.123632891797,
               -0.171503832236, 0.190899511506, -0.171503832236, 0.123632891797, -0.069975429105,
               0.029662754736, -0.008520738659, 0.001254597714])
DFB = np.array([-0.057717512738, 0.199278374994, -0.292668277650, 0.244537361546, -0.134605018019,
                0.056184263460, -0.015009191593])
DFB_F32 = DFB.astype(np.float32).astype(np.float64)


def _golden_case(g, nk=5):
    imax, jmax = int(g["imax"]), int(g["jmax"])
    c = config.random_case(imax, jmax, nk=nk, D=0, seed=0, mesh_random=False)
    c.dm, c.dn, c.gamma, c.Tref = float(g["dm"]), float(g["dn"]), float(g["gamma"]), float(g["Tref"])
    c.rhob, c.psib = g["rhob"], g["psib"]
    return c


# ---------------------------------------------------------------------------
# against the reference's compiled routines
# ---------------------------------------------------------------------------
def test_golden_vectors_derivatives_transport_ccpml():
    g = np.load(GOLDEN)
    c = _golden_case(g)
    o = Oracle(c)
    assert np.array_equal(o.deropm(g["U"]), g["dUdm"])      # D=0: no metric region
    assert np.array_equal(o.deropn(g["U"]), g["dUdn"])
    mi, kk = o.transport(g["T"])
    assert np.array_equal(mi, g["mi"]) and np.array_equal(kk, g["k"])
    assert np.array_equal(o.ccpml(g["U"]), g["U_ccpml"])
    coef = dict(zip(g["coef_names"], g["coef_values"]))
    ref = c.transport()
    for key, val in coef.items():
        assert ref[str(key)] == val, key


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("nk,D,seed", [(5, 6, 1), (5, 9, 2)])
def test_oracle_matches_reference_routines(nk, D, seed):
    c = config.random_case(47, 38, nk=nk, D=D, seed=seed)
    o, r = Oracle(c), Reference(c.imax, c.jmax, c.dm, c.dn, c.gamma, c.Tref)
    U = c.U0
    # metricsX / metricsY multiply by meshx(i) / meshy(j) in the buffer layers only
    metx = np.where((np.arange(1, c.imax + 1) <= D) | (np.arange(1, c.imax + 1) >= c.imax - D + 1), c.meshx, 1.0)
    mety = np.where((np.arange(1, c.jmax + 1) <= D) | (np.arange(1, c.jmax + 1) >= c.jmax - D + 1), c.meshy, 1.0)
    assert np.array_equal(o.deropm(U), r.deropm(U) * metx[None, :, None])
    assert np.array_equal(o.deropn(U), r.deropn(U) * mety[:, None, None])
    T = c.gamma * U[..., 3] / U[..., 0]
    mo, ko = o.transport(T)
    mr, kr = r.transport(T)
    assert np.array_equal(mo, mr) and np.array_equal(ko, kr)
    assert np.array_equal(o.ccpml(U), r.ccpml(U, c.rhob, c.psib))


# ---------------------------------------------------------------------------
# closed forms: DRP derivative
# ---------------------------------------------------------------------------
def _plain_case(imax, jmax, h=1.0):
    c = config.random_case(imax, jmax, nk=4, D=0, seed=0, mesh_random=False)
    c.dm = c.dn = h
    return c


def test_drp_linear_and_wavenumber():
    c = _plain_case(64, 20, h=0.5)
    o = Oracle(c)
    x = np.arange(c.imax) * c.dm
    f = np.broadcast_to(3.0 * x + 1.0, (c.jmax, c.imax)).copy()
    d = o.deropm(f)
    # interior Tam-Webb stencil: 2*(a1 + 2 a2 + 3 a3) = 1 - 2.2e-12
    assert np.allclose(d[:, 3:-3], 3.0, rtol=1e-10, atol=0)
    # one-sided stencils: first moment 1 to ~4e-9 (7-digit tables)
    assert np.allclose(d, 3.0, rtol=2e-8, atol=0)
    assert not np.allclose(d, 3.0, rtol=1e-12, atol=0)
    # DRP: modified wavenumber k* dx = 2 sum a_s sin(s k dx)
    a = np.array([0.770882380518, -0.166705904415, 0.0208431427703])
    for kdx in (0.3, 0.9, 1.5):
        f = np.sin(kdx * x / c.dm)
        g = np.broadcast_to(f, (c.jmax, c.imax)).copy()
        d = o.deropm(g)[0, 10:-10]
        kstar = 2.0 * sum(a[s] * np.sin((s + 1) * kdx) for s in range(3)) / c.dm
        assert np.allclose(d, kstar * np.cos(kdx * x[10:-10] / c.dm), atol=1e-12)
    # x and y operators agree under transposition
    rng = np.random.default_rng(3)
    sq = _plain_case(30, 30)
    osq = Oracle(sq)
    u = rng.standard_normal((30, 30))
    assert np.array_equal(osq.deropm(u), osq.deropn(u.T).T)


def test_drp_fourth_order_convergence():
    errs = []
    for N in (40, 80, 160):
        h = 2 * np.pi / (N - 1)
        c = _plain_case(N, 16, h=h)
        x = np.arange(N) * h
        u = np.broadcast_to(np.sin(1.3 * x) + np.cos(0.7 * x), (16, N)).copy()
        exact = 1.3 * np.cos(1.3 * x) - 0.7 * np.sin(0.7 * x)
        errs.append(np.abs(Oracle(c).deropm(u)[0] - exact).max())
    assert errs[0] / errs[1] > 12 and errs[1] / errs[2] > 12


# ---------------------------------------------------------------------------
# filter: independent triangular-solve formulation of the in-place sweeps
# ---------------------------------------------------------------------------
def _sweep_block(u, rows, weights, sigma):
    """Gauss-Seidel block `rows` (in sweep order) as (I + sigma L) x = b."""
    pos = {q: a for a, q in enumerate(rows)}
    nb = len(rows)
    M = np.eye(nb)
    b = np.empty(nb)
    for a, q in enumerate(rows):
        acc = 0.0
        for qq, w in weights(q):
            if qq in pos and pos[qq] < a:
                M[a, pos[qq]] += sigma * w
            else:
                acc += w * u[qq]
        b[a] = u[q] - sigma * acc
    x = solve_triangular(M, b, lower=True)
    out = u.copy()
    out[rows] = x
    return out


def filter_line_matrix(u, sigma, dfb):
    n = u.shape[0]
    u = _sweep_block(u, list(range(1, 7)), lambda q: [(q + s, dfb[s + 1]) for s in range(-1, 6)], sigma)
    u = _sweep_block(u, list(range(6, n - 6)), lambda q: [(q + s, DF[s + 6]) for s in range(-6, 7)], sigma)
    u = _sweep_block(u, list(range(n - 2, n - 8, -1)), lambda q: [(q - s, dfb[s + 1]) for s in range(-1, 6)],
                     sigma)
    return u


def test_filter_coefficients_closed_forms():
    assert abs(DF.sum()) < 1e-15                      # constants pass through
    s = np.arange(-6, 7)
    assert abs(np.sum(DF * np.cos(np.pi * s)) - 1.0) < 1e-4   # full damping at k dx = pi
    assert np.allclose(DF, DF[::-1])                  # centred
    assert abs(DFB.sum()) < 1e-15
    # float32 literals of filtering.f90:114-120 no longer sum to zero
    assert 5e-9 < abs(DFB_F32.sum()) < 2e-8


@pytest.mark.parametrize("n,sigma", [(16, 0.2), (41, 0.1), (64, 0.2)])
def test_filter_matches_triangular_solve(n, sigma):
    c = _plain_case(n, n)
    c.sigmaf = sigma
    o = Oracle(c)
    rng = np.random.default_rng(n)
    U = rng.standard_normal((n, n, 4))
    fx = o.filterx(U)
    fy = o.filtery(U)
    for j in (0, 3, n - 1):
        for k in range(4):
            assert rel_max(fx[j, :, k], filter_line_matrix(U[j, :, k], sigma, DFB)) < 1e-14
    for i in (0, 5, n - 2):
        for k in range(4):
            assert rel_max(fy[:, i, k], filter_line_matrix(U[:, i, k], sigma, DFB_F32)) < 1e-14


def test_filter_constant_state_and_f32_quirk():
    c = _plain_case(64, 64)
    o = Oracle(c)
    U = np.full((64, 64, 4), 2.5)
    assert np.abs(o.filterx(U) - 2.5).max() < 1e-15
    fy = o.filtery(U)
    # rows j=2..7 and jmax-6..jmax-1 move by sigma*|sum dfb_f32|*c per sweep
    shift = np.abs(fy[:, 10, 0] - 2.5)
    expect = c.sigmaf * abs(DFB_F32.sum()) * 2.5
    assert 0.3 * expect < shift[1] < 3 * expect
    # the interior recurrence carries that drift inward, decaying ~0.33 per point
    assert np.abs(fy[24:-24] - 2.5).max() < 1e-15
    assert shift[12] < shift[8] < shift[6]


# ---------------------------------------------------------------------------
# LDDRK6 (rk.f90:27-46, :153-169)
# ---------------------------------------------------------------------------
def _rk_case():
    c = _plain_case(16, 16)
    c.sigmaf = 0.0      # filter becomes the identity
    c.dt = 1.0
    return c


def test_lddrk_amplification_polynomial_is_fourth_order():
    c = _rk_case()
    o = Oracle(c)
    U0 = np.ones((16, 16, 4))
    zs = np.linspace(-1.2, 0.4, 13)
    amp = []
    for z in zs:
        out = o.rk_euler_cb(U0, 0, lambda U, irk, t, z=z: z * U)
        amp.append(out[8, 8, 0])
    coef = np.polynomial.polynomial.polyfit(zs, amp, 6)
    assert np.allclose(coef[:5], [1, 1, 1 / 2, 1 / 6, 1 / 24], atol=1e-9)
    assert coef[5] > 0 and coef[6] > 0


def test_lddrk_stage_times_integrate_cubics():
    c = _rk_case()
    c.dt = 0.37
    o = Oracle(c)
    U0 = np.zeros((16, 16, 4))
    for mexp in range(4):
        out = o.rk_euler_cb(U0, 0, lambda U, irk, t, mexp=mexp: np.full_like(U, t ** mexp))
        # the tables carry 12 digits: quadrature exact to ~1e-12 relative
        exact = c.dt ** (mexp + 1) / (mexp + 1)
        assert abs(out[5, 5, 2] - exact) < 1e-11 * exact


# ---------------------------------------------------------------------------
# RHS: manufactured solution (primitive-variable Navier-Stokes)
# ---------------------------------------------------------------------------
def _spectral_fields(N, L, nk=4):
    x = np.arange(N) * L / N
    X, Y = np.meshgrid(x, x)
    kx = 2 * np.pi / L
    rho = 1.0 + 0.1 * np.sin(kx * X) * np.cos(kx * Y)
    u = 0.3 + 0.05 * np.cos(kx * X + 0.3) * np.sin(2 * kx * Y)
    v = -0.1 + 0.04 * np.sin(2 * kx * X) * np.sin(kx * Y + 0.2)
    p = 1 / 1.4 + 0.03 * np.cos(kx * X - 0.4) * np.cos(kx * Y)
    fields = [rho, u, v, p]
    if nk >= 5:
        fields.append(0.5 + 0.2 * np.sin(kx * X + 0.7) * np.cos(2 * kx * Y - 0.1))
    return X, Y, np.stack(fields, axis=-1)


def _dx(f, L, axis):
    N = f.shape[axis]
    k = 2j * np.pi * np.fft.fftfreq(N, d=L / N)
    shape = [1, 1]
    shape[axis] = N
    return np.real(np.fft.ifft(np.fft.fft(f, axis=axis) * k.reshape(shape), axis=axis))


def _pde_rhs(Uf, L, Re, Pr, tc, gamma=1.4, Tref=273.15, Pe=1000.0, alpha_cond=0.5):
    rho, u, v, p = (Uf[..., k] for k in range(4))
    dx = lambda f: _dx(f, L, 1)  # noqa: E731
    dy = lambda f: _dx(f, L, 0)  # noqa: E731
    T = gamma * p / rho
    tt = Tref * T
    mu = np.exp(tc["A_mi"] * np.log(tt) + tc["B_mi"] / tt + tc["C_mi"] / tt ** 2 + tc["D_mi"]) / tc["MI0"]
    kk = np.exp(tc["A_k"] * np.log(tt) + tc["B_k"] / tt + tc["C_k"] / tt ** 2 + tc["D_k"]) / tc["K0"]
    ux, uy, vx, vy = dx(u), dy(u), dx(v), dy(v)
    txx = mu * (2 * ux - 2 / 3 * (ux + vy))
    tyy = mu * (2 * vy - 2 / 3 * (ux + vy))
    txy = mu * (uy + vx)
    phi = txx * ux + txy * (uy + vx) + tyy * vy
    out = np.empty_like(Uf)
    out[..., 0] = -(u * dx(rho) + v * dy(rho) + rho * (ux + vy))
    out[..., 1] = -(u * ux + v * uy + dx(p) / rho) + (dx(txx) + dy(txy)) / (Re * rho)
    out[..., 2] = -(u * vx + v * vy + dy(p) / rho) + (dx(txy) + dy(tyy)) / (Re * rho)
    out[..., 3] = (-(u * dx(p) + v * dy(p) + gamma * p * (ux + vy)) + (gamma - 1) / Re * phi
                   + (dx(kk * dx(T)) + dy(kk * dy(T))) / (Re * Pr))
    if Uf.shape[-1] >= 5:
        # passive scalar: derivs.f90:36-43 forms the flux rho*alpha*(psi_x + psi_y)
        # (the sum of both derivatives) and differentiates it in both
        # directions; eqeuler.f90:94-95 divides by Pe*rho
        psi = Uf[..., 4]
        px, py = dx(psi), dy(psi)
        F = rho * alpha_cond * (px + py)
        out[..., 4] = -(u * px + v * py - (dx(F) + dy(F)) / (Pe * rho))
    return out


def test_rhs_converges_to_navier_stokes():
    L = 10.0
    errs = []
    for N in (48, 96):
        X, Y, Uf = _spectral_fields(N, L)
        c = _plain_case(N, N, h=L / N)
        c.Re, c.Pr = 5.0, 0.7
        exact = _pde_rhs(Uf, L, c.Re, c.Pr, c.transport())
        got = Oracle(c).rhs(Uf)
        inner = slice(N // 4, 3 * N // 4)
        errs.append(max(rel_max(got[inner, inner, k], exact[inner, inner, k]) for k in range(4)))
    assert errs[1] < 2e-4 and errs[0] / errs[1] > 8, errs


def _scalar_case(N, h):
    c = config.random_case(N, N, nk=5, D=0, seed=0, mesh_random=False)
    c.dm = c.dn = h
    c.Re, c.Pr = 5.0, 0.7
    # Pe of order one so that the diffusion term is as large as the
    # advection term (the reference's Pe = 1000 would hide it)
    c.Pe, c.alpha_cond = 0.8, 0.5
    return c


def test_scalar_rhs_converges_to_advection_diffusion():
    """The nk=5 passive-scalar equation (derivs.f90:36-43, eqeuler.f90:94-95)
    at the PDE level: psi_t = -(u psi_x + v psi_y - (d/dx + d/dy)[rho alpha
    (psi_x + psi_y)] / (Pe rho)); the oracle's fourth-order discretisation
    converges to it on a periodic manufactured field."""
    L = 10.0
    errs = []
    for N in (48, 96):
        X, Y, Uf = _spectral_fields(N, L, nk=5)
        c = _scalar_case(N, L / N)
        exact = _pde_rhs(Uf, L, c.Re, c.Pr, c.transport(), Pe=c.Pe, alpha_cond=c.alpha_cond)
        got = Oracle(c).rhs(Uf)
        inner = slice(N // 4, 3 * N // 4)
        errs.append(rel_max(got[inner, inner, 4], exact[inner, inner, 4]))
        # components 1-4 do not see psi
        four = Oracle(c.with_nk(4)).rhs(Uf[..., :4])
        assert np.array_equal(got[..., :4], four)
    assert errs[1] < 2e-4 and errs[0] / errs[1] > 8, errs


def test_scalar_pin_rejects_plausible_misreadings():
    """The scalar pin separates the reference's form from nearby misreadings:
    each of them differs from the oracle by far more than its discretisation
    error."""
    L, N = 10.0, 96
    X, Y, Uf = _spectral_fields(N, L, nk=5)
    c = _scalar_case(N, L / N)
    got = Oracle(c).rhs(Uf)[..., 4]
    inner = slice(N // 4, 3 * N // 4)
    rho, u, v, psi = Uf[..., 0], Uf[..., 1], Uf[..., 2], Uf[..., 4]
    dx = lambda f: _dx(f, L, 1)  # noqa: E731
    dy = lambda f: _dx(f, L, 0)  # noqa: E731
    px, py = dx(psi), dy(psi)
    a, Pe = c.alpha_cond, c.Pe
    variants = {
        # Fickian flux: each direction differentiated on its own
        "fick": -(u * px + v * py - (dx(rho * a * px) + dy(rho * a * py)) / (Pe * rho)),
        # no 1/rho on the diffusion term
        "no_inv_rho": -(u * px + v * py - (dx(rho * a * (px + py)) + dy(rho * a * (px + py))) / Pe),
        # diffusion sign flipped
        "sign": -(u * px + v * py + (dx(rho * a * (px + py)) + dy(rho * a * (px + py))) / (Pe * rho)),
    }
    for name, f in variants.items():
        assert rel_max(got[inner, inner], f[inner, inner]) > 1e-2, name


def test_metric_everywhere_closed_form():
    """The oracle's metric_everywhere switch (an extension of this build for
    stretched grids, used by the mixing layer) multiplies every derivative by
    meshx(i) / meshy(j) exactly once, at every point: bitwise the unit-metric
    derivative times the profile.  Without it, interior points stay
    unmultiplied (the reference's region tests, diff.f90:276-350)."""
    import copy
    c = config.random_case(40, 36, nk=4, D=6, seed=9)   # random meshx, meshy
    c.metric_everywhere = True
    flat = copy.copy(c)
    flat.meshx = np.ones_like(c.meshx)
    flat.meshy = np.ones_like(c.meshy)
    U = c.U0
    o, of = Oracle(c), Oracle(flat)
    assert np.array_equal(o.deropm(U), of.deropm(U) * c.meshx[None, :, None])
    assert np.array_equal(o.deropn(U), of.deropn(U) * c.meshy[:, None, None])
    ref = copy.copy(c)
    ref.metric_everywhere = False
    D = c.D
    inner = (slice(D + 1, -D - 1), slice(D + 1, -D - 1))
    assert np.array_equal(Oracle(ref).deropm(U)[inner], of.deropm(U)[inner])
    assert not np.array_equal(o.deropm(U)[inner], of.deropm(U)[inner])


def test_rhs_uniform_state_is_steady():
    c = _plain_case(32, 24)
    U = np.empty((24, 32, 4))
    U[...] = [1.0, 0.4, -0.2, 1 / 1.4]
    assert np.abs(Oracle(c).rhs(U)).max() < 1e-13


# ---------------------------------------------------------------------------
# buffer zone, boundary conditions, source
# ---------------------------------------------------------------------------
def test_buffer_closed_forms():
    c = config.random_case(40, 36, nk=4, D=7, seed=4)
    rng = np.random.default_rng(1)
    c.sigmax = rng.random(40)
    c.sigmay = rng.random(36)
    o = Oracle(c)
    U = c.U0
    out = o.bufferxy(U)
    Ub = c.Ub
    D = c.D
    # interior untouched, corners get both layers
    assert np.array_equal(out[D:-D, D:-D], U[D:-D, D:-D])
    corner = Ub[2, 3] + c.sigmax[3] * c.sigmay[2] * (U[2, 3] - Ub[2, 3])
    assert rel_max(out[2, 3], corner) < 1e-14
    c.sigmax[:] = 1.0
    c.sigmay[:] = 1.0
    assert np.allclose(Oracle(c).bufferxy(U), U, rtol=0, atol=1e-15)
    c.sigmax[:] = 0.0
    out0 = Oracle(c).bufferxy(U)
    assert np.array_equal(out0[:, :D], Ub[:, :D])


def test_acoustic_source_closed_form():
    c = config.random_case(41, 41, nk=4, D=0, seed=0, mesh_random=False)
    c.amplitude, c.wf, c.r0_acous = 0.8, 1.3, 3.0
    c.m_pt_acous, c.n_pt_acous = float(c.m[20]), float(c.n[20])
    s = Oracle(c).acousticfont(0.25, 7)
    peak = 0.8 * np.sin(1.3 * ((7 + 0.25) * c.dt))
    assert abs(s[20, 20] - peak) < 1e-15
    r2 = (c.m[25] - c.m[20]) ** 2 + (c.n[18] - c.n[20]) ** 2
    assert abs(s[18, 25] / peak - np.exp(-np.log(2) * r2 / 9.0)) < 1e-14


def test_full_step_base_state_invariance():
    """A state equal to the quiescent base flow stays at it up to the f32
    y-filter literals' drift (about sigma*8.4e-9 of the state near j edges)."""
    c = config.acoustic_pulse_case(101, D=6)
    c.U0 = c.Ub.copy()
    out = Oracle(c).rk_euler(c.U0, 0)
    dev = np.abs(out - c.Ub).max(axis=(1, 2))
    assert dev[30:-30].max() < 1e-15
    assert 1e-9 < dev.max() < 1e-7
    c.sigmaf = 0.0   # without the filter the base state is steady to rounding
    assert np.abs(Oracle(c).rk_euler(c.U0, 0) - c.Ub).max() < 1e-15


def test_nk5_scalar_does_not_feed_back():
    c5 = config.random_case(40, 30, nk=5, D=5, seed=3)
    c4 = c5.with_nk(4)
    o5, o4 = Oracle(c5), Oracle(c4)
    a = o5.rk_euler(c5.U0, 0)
    b = o4.rk_euler(c4.U0, 0)
    assert np.array_equal(a[..., :4], b)


def test_openmp_build_is_bitwise_serial():
    c = config.random_case(50, 44, nk=5, D=6, seed=7)
    a = Oracle(c).run(c.U0, 0, 2)
    b = Oracle(c, parallel=True).run(c.U0, 0, 2)
    assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# periodic x (no reference counterpart: the build's extension for the
# temporal mixing layer)
# ---------------------------------------------------------------------------
def _periodic_case(imax=96, jmax=40, nk=4, seed=5, D=6):
    c = config.random_case(imax, jmax, nk=nk, D=D, seed=seed, mesh_random=True)
    c.periodic_x = True
    return c


def filter_ring_matrix(u, sigma):
    """The ring sweep as one linear solve: with D the circulant 13-point
    filter and L its part on points the sweep has already updated (column
    (q+s) mod n < q), the sweep is (I + sigma L) u' = (I - sigma (D - L)) u,
    L strictly lower triangular."""
    n = u.shape[0]
    Dm = np.zeros((n, n))
    for q in range(n):
        for s in range(-6, 7):
            Dm[q, (q + s) % n] += DF[s + 6]
    L = np.tril(Dm, -1)
    rhs = u - sigma * ((Dm - L) @ u)
    return solve_triangular(np.eye(n) + sigma * L, rhs, lower=True)


def test_periodic_filter_against_triangular_solve_and_fourier_mode():
    c = _periodic_case(96, 20)
    o = Oracle(c)
    rng = np.random.default_rng(8)
    U = rng.standard_normal((20, 96, 4))
    got = o.filterx(U)
    for j in (0, 7, 19):
        ref = filter_ring_matrix(U[j, :, 2], c.sigmaf)
        assert rel_max(got[j, :, 2], ref) < 1e-14
    # the y filter is untouched by the x periodicity
    c2 = _periodic_case(96, 20)
    c2.periodic_x = False
    assert np.array_equal(o.filtery(U), Oracle(c2).filtery(U))
    # a Fourier mode away from the seam: u' = G u with the in-place sweep's
    # amplification G = (1 - sigma (d0 + A+)) / (1 + sigma A-), A+- the
    # taps ahead of / behind the point
    n = 200
    c3 = _periodic_case(n, 16)
    s = np.arange(-6, 7)
    for m in (7, 31, 70):
        k = 2 * np.pi * m / n
        x = np.arange(n)
        Uk = np.zeros((16, n, 4))
        Uk[..., 1] = np.cos(k * x + 0.3)[None, :]
        out = Oracle(c3).filterx(Uk)[5, :, 1]
        ap = np.sum(DF[s > 0] * np.exp(1j * k * s[s > 0]))
        am = np.sum(DF[s < 0] * np.exp(1j * k * s[s < 0]))
        G = (1 - c3.sigmaf * (DF[6] + ap)) / (1 + c3.sigmaf * am)
        exact = np.real(G * np.exp(1j * (k * x + 0.3)))
        inner = slice(60, n - 7)     # the seam's transient has decayed (0.33^60)
        assert np.abs(out[inner] - exact[inner]).max() < 1e-14, m
        assert np.abs(out[:6] - exact[:6]).max() > 1e-11    # the seam is visible


def test_periodic_derivative_is_the_wrapped_drp():
    n = 64
    c = _periodic_case(n, 12)
    c.meshx = np.ones(n)
    c.dm = 0.5
    o = Oracle(c)
    a = np.array([0.770882380518, -0.166705904415, 0.0208431427703])
    for m in (1, 5, 13):
        kdx = 2 * np.pi * m / n
        x = np.arange(n)
        f = np.broadcast_to(np.sin(kdx * x), (12, n)).copy()
        kstar = 2.0 * sum(a[t] * np.sin((t + 1) * kdx) for t in range(3)) / c.dm
        assert np.allclose(o.deropm(f), kstar * np.cos(kdx * x)[None, :], rtol=0, atol=1e-13)
    # no x metric layers: meshx is ignored unless applied everywhere
    c.meshx = 1.0 + np.random.default_rng(2).random(n)
    assert np.array_equal(Oracle(c).deropm(f), o.deropm(f))


def test_periodic_pointwise_operators():
    c = _periodic_case(50, 36, nk=5, D=7)
    c.sigmax = np.random.default_rng(3).random(50)
    o = Oracle(c)
    U = c.U0
    # no inflow/outflow columns; rows j = 1, jmax on every column (cc.f90:38-59)
    cc = o.ccpml(U)
    assert np.array_equal(cc[1:-1], U[1:-1])
    assert np.array_equal(cc[0, :, :3], 2.0 * U[1, :, :3] - U[2, :, :3])
    assert np.all(cc[0, :, 3] == 1.0 / c.gamma) and np.all(cc[-1, :, 4] == 1.0)
    # no x buffer layers: the y layers alone
    b = o.bufferxy(U)
    D = c.D
    assert np.array_equal(b[D:-D], U[D:-D])
    # the LDDRK update reaches every column
    rng = np.random.default_rng(4)
    dudt, w = rng.standard_normal(U.shape), rng.standard_normal(U.shape)
    Un, _ = o.rk_steep(-0.7, 0.8, U, dudt, w)
    assert not np.array_equal(Un[1:-1, 0], U[1:-1, 0]) and not np.array_equal(Un[1:-1, -1], U[1:-1, -1])
    assert np.array_equal(Un[0], U[0]) and np.array_equal(Un[-1], U[-1])


def test_periodic_step_commutes_with_x_shifts():
    """Without the filter's seam (sigma_f = 0) every x operation of the
    periodic step is a wrapped stencil, the same arithmetic at every column:
    the step commutes with rolling the state in x, bitwise."""
    c = _periodic_case(60, 48, nk=5, D=8)
    c.sigmaf = 0.0
    c.meshx = np.ones(60)
    o = Oracle(c)
    a = o.rk_euler(c.U0, 0)
    for r in (1, 17):
        cr = _periodic_case(60, 48, nk=5, D=8)
        cr.sigmaf = 0.0
        cr.meshx = np.ones(60)
        cr.U0 = np.roll(c.U0, r, axis=1)
        b = Oracle(cr).rk_euler(cr.U0, 0)
        assert np.array_equal(np.roll(a, r, axis=1), b)
