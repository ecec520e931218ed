"""Host-side set-up: parameter blocks, grids, buffer profiles, base flows and
initial conditions.

`grid_profiles` and `mixing_layer_fields` restate the reference's one-time
set-up (init.f90:56-206 init_dados, :350-404 init_variables) so that a
reference-style run can be reproduced; `read_dados` reads the reference's
namelist file (dados.in, init.f90:18-47).  The workload builders below
(`acoustic_pulse_case`, `pulse_vortex_case`, ...) construct the benchmark
configurations around the same kernels.

Arrays use the reference's memory order: a state U(1:nk,1:imax,1:jmax) is a
C-contiguous numpy array of shape (jmax, imax, nk).
"""
from __future__ import annotations

import dataclasses
import math
import re
from pathlib import Path
from typing import Optional

import numpy as np

GAMMA = 1.4          # init.f90:361
PE = 1000.0          # global.f90:38
ALPHA_COND = 0.5     # global.f90:39


def transport_coefficients(Tref: float) -> dict:
    """N2 fits and reference values (coefficients.f90:71-123)."""
    c = dict(A_mi=0.62526577, B_mi=-31.779652, C_mi=-1640.7913, D_mi=1.7454992,
             A_k=0.85372829, B_k=105.18665, C_k=-12229.753, D_k=0.48299104)
    T = Tref
    c["K0"] = math.exp(c["A_k"] * math.log(1.0 * T) + c["B_k"] / T + c["C_k"] / T / T + c["D_k"])
    c["MI0"] = math.exp(c["A_mi"] * math.log(1.0 * T) + c["B_mi"] / T + c["C_mi"] / T / T + c["D_mi"])
    return c


@dataclasses.dataclass
class Case:
    """Everything the reference's `module global` holds for the hot path."""
    name: str
    imax: int
    jmax: int
    nk: int
    D: int
    dm: float
    dn: float
    dt: float
    sigmaf: float
    m: np.ndarray
    n: np.ndarray
    meshx: np.ndarray
    meshy: np.ndarray
    sigmax: np.ndarray
    sigmay: np.ndarray
    rhob: np.ndarray
    psib: np.ndarray
    Ub: np.ndarray          # (jmax, imax, nk)
    U0: np.ndarray          # (jmax, imax, nk)
    gamma: float = GAMMA
    Re: float = 1.0e4
    Pr: float = 0.71
    Pe: float = PE
    alpha_cond: float = ALPHA_COND
    Tref: float = 273.15
    amplitude: float = 0.0
    wf: float = 1.0
    r0_acous: float = 0.03
    m_pt_acous: float = 0.0
    n_pt_acous: float = 0.0
    inviscid: bool = False
    metric_everywhere: bool = False
    periodic_x: bool = False   # no reference counterpart: periodic in x (NS_FLAG_PERIODIC_X)

    @property
    def npoints(self) -> int:
        return self.imax * self.jmax

    def transport(self) -> dict:
        return transport_coefficients(self.Tref)

    def with_nk(self, nk: int) -> "Case":
        """Same case restricted to the first nk components (nk <= self.nk)."""
        assert nk <= self.nk
        return dataclasses.replace(self, nk=nk, Ub=np.ascontiguousarray(self.Ub[..., :nk]),
                                   U0=np.ascontiguousarray(self.U0[..., :nk]))


# ---------------------------------------------------------------------------
# init.f90:56-206 (init_dados)
# ---------------------------------------------------------------------------
def grid_profiles(imax: int, jmax: int, mmin: float, mmax: float, nmin: float, nmax: float,
                  D: int, c1b: float, c2b: float, sponge_left_bottom: bool = False) -> dict:
    """Grid coordinates, stretch factors and buffer profiles.

    The reference tests `i <= imaxpml` / `j <= jmaxpml` (init.f90:168,192) on
    two module variables that are never assigned, so the left and bottom
    profiles stay 0 (the buffer there resets to the base flow every stage).
    `sponge_left_bottom=True` applies the right/top profile mirrored instead.
    """
    f32 = np.float32
    dm = (mmax - mmin) / float(f32(imax - 1))   # real(imax-1) is default real
    dn = (nmax - nmin) / float(f32(jmax - 1))
    Dx, Dy = dm * D, dn * D
    xli, xld = mmin + Dx, mmax - Dx
    yli, yls = nmin + Dy, nmax - Dy
    m = np.array([mmin + float(f32(i - 1)) * dm for i in range(1, imax + 1)])
    n = np.array([nmin + float(f32(j - 1)) * dn for j in range(1, jmax + 1)])
    Aa = 2.0  # init.f90:119-120 (alpham = 2)

    def stretch(c, c0, Dl):
        a = abs((c - c0) / Dl)
        return 1.0 / (1.0 + Aa * (a * a))

    meshx = np.ones(imax)
    meshy = np.ones(jmax)
    for i in range(1, imax + 1 if D > 0 else 1):
        if 1 <= i <= D + 1:
            meshx[i - 1] = stretch(m[i - 1], xli, Dx)
        if imax - D <= i <= imax:
            meshx[i - 1] = stretch(m[i - 1], xld, Dx)
    for j in range(1, jmax + 1 if D > 0 else 1):
        if 1 <= j <= D + 1:
            meshy[j - 1] = stretch(n[j - 1], yli, Dy)
        if jmax - D + 1 <= j <= jmax:
            meshy[j - 1] = stretch(n[j - 1], yls, Dy)

    def prof(x):
        return (1.0 - c1b * x * x) * (1.0 - (1.0 - math.exp(c2b * x * x)) / (1.0 - math.exp(c2b)))

    imaxpml = D if sponge_left_bottom else 0
    sigmax = np.zeros(imax)
    for i in range(1, imax + 1):
        if 1 <= i <= imaxpml:
            sigmax[i - 1] = prof((m[i - 1] - xli) / (mmin - xli))
        if imax - D + 1 <= i <= imax:
            sigmax[i - 1] = prof((m[i - 1] - xld) / (mmax - xld))
    sigmay = np.zeros(jmax)
    for j in range(1, jmax + 1):
        if 1 <= j <= imaxpml:
            sigmay[j - 1] = prof((n[j - 1] - yli) / (nmin - yli))
        if jmax - D + 1 <= j <= jmax:
            sigmay[j - 1] = prof((n[j - 1] - yls) / (nmax - yls))
    return dict(dm=dm, dn=dn, m=m, n=n, meshx=meshx, meshy=meshy, sigmax=sigmax, sigmay=sigmay)


# ---------------------------------------------------------------------------
# init.f90:350-404 (init_variables): tanh mixing layer
# ---------------------------------------------------------------------------
def mixing_layer_fields(n: np.ndarray, imax: int, nk: int, U1: float, U2: float, T1: float,
                        T2: float, gamma: float = GAMMA, delta: float = 0.4):
    psi1, psi2 = 1.0, 0.0
    uba = 0.5 * (U1 + U2 + (U1 - U2) * np.tanh(2.0 * n / delta))
    Tb = (T1 * (uba - U2) / (U1 - U2) + T2 * (U1 - uba) / (U1 - U2)
          + (gamma - 1.0) / 2.0 * (U1 - uba) * (uba - U2))
    rhob = 1.0 / Tb
    psib = 0.5 * (psi1 + psi2 + (psi1 - psi2) * np.tanh(2.0 * n / delta))
    jmax = n.shape[0]
    Ub = np.zeros((jmax, imax, nk))
    Ub[:, :, 0] = rhob[:, None]
    Ub[:, :, 1] = uba[:, None]
    Ub[:, :, 3] = 1.0 / gamma
    if nk >= 5:
        Ub[:, :, 4] = psib[:, None]
    return rhob, psib, Ub


def read_dados(path) -> dict:
    """Parse the reference's namelist file (dados.in; groups read at
    init.f90:28-47).  Returns a flat dict of lower-case names."""
    text = Path(path).read_text()
    out: dict = {}
    for line in text.splitlines():
        line = line.split("!")[0].strip()
        mm = re.match(r"^([A-Za-z_][A-Za-z0-9_]*)\s*=\s*([^,]+)", line)
        if not mm:
            continue
        key, val = mm.group(1).lower(), mm.group(2).strip().lower().replace("d", "e")
        try:
            out[key] = int(val)
        except ValueError:
            out[key] = float(val)
    return out


def reference_case(dados: Optional[dict] = None, nk: int = 5) -> Case:
    """The reference's own run: dados.in defaults + mixing-layer base flow."""
    d = dict(imax=521, jmax=481, dt=0.01, sigmaf=0.1, r0_acous=0.03, amplitude=0.8,
             m_pt_acous=0.0, n_pt_acous=0.0, wf=1.0, mmin=-6.5, mmax=19.5, nmin=-2.4,
             nmax=2.4, u1=0.9, u2=0.1, re=10000.0, pr=0.71, t1=1.0, t2=0.5, tref=273.15,
             d=90, c1b=0.01, c2b=20.0)
    if dados:
        d.update(dados)
    g = grid_profiles(d["imax"], d["jmax"], d["mmin"], d["mmax"], d["nmin"], d["nmax"],
                      int(d["d"]), d["c1b"], d["c2b"])
    rhob, psib, Ub = mixing_layer_fields(g["n"], d["imax"], nk, d["u1"], d["u2"], d["t1"], d["t2"])
    return Case(name="reference-mixing-layer", imax=d["imax"], jmax=d["jmax"], nk=nk, D=int(d["d"]),
                dm=g["dm"], dn=g["dn"], dt=d["dt"], sigmaf=d["sigmaf"], m=g["m"], n=g["n"],
                meshx=g["meshx"], meshy=g["meshy"], sigmax=g["sigmax"], sigmay=g["sigmay"],
                rhob=rhob, psib=psib, Ub=Ub, U0=Ub.copy(), Re=d["re"], Pr=d["pr"], Tref=d["tref"],
                amplitude=d["amplitude"], wf=d["wf"], r0_acous=d["r0_acous"],
                m_pt_acous=d["m_pt_acous"], n_pt_acous=d["n_pt_acous"])


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def _uniform_case(name, N, half, D, u0, sigmaf, nk, Re, Pr, inviscid, cfl=0.5, M=None):
    g = grid_profiles(N, N, -half, half, -half, half, D, 0.01, 20.0)
    p0 = 1.0 / GAMMA
    Ub = np.zeros((N, N, nk))
    Ub[..., 0] = 1.0
    Ub[..., 1] = u0
    Ub[..., 3] = p0
    rhob = np.ones(N)
    psib = np.zeros(N)
    dt = cfl * min(g["dm"], g["dn"]) / (1.0 + abs(u0))
    return g, Ub, rhob, psib, dt, p0


def _add_pulse(U, X, Y, xc, yc, amp, b, gamma=GAMMA):
    """Isentropic Gaussian pressure pulse p' = amp*exp(-ln2 r^2/b^2) on top of U."""
    p_old = U[..., 3].copy()
    r2 = (X - xc) ** 2 + (Y - yc) ** 2
    U[..., 3] = p_old + amp * np.exp(-math.log(2.0) * r2 / (b * b))
    U[..., 0] = U[..., 0] * (U[..., 3] / p_old) ** (1.0 / gamma)


def acoustic_pulse_case(N: int = 201, D: int = 30, steps_cfl: float = 0.5, sigmaf: float = 0.2,
                        nk: int = 4, amp_rel: float = 1e-3, half_width: float = 3.0,
                        u0: float = 0.0) -> Case:
    """2-D Euler, uniform grid dx=dy=1 centred on 0, quiescent (or uniform u0)
    medium rho0=1, p0=1/gamma; Gaussian pulse of amplitude amp_rel*p0 at the
    centre with matching isentropic density; FP64; CFL 0.5."""
    half = (N - 1) / 2.0
    g, Ub, rhob, psib, dt, p0 = _uniform_case("pulse", N, half, D, u0, sigmaf, nk, 1e300, 0.71, True,
                                               cfl=steps_cfl)
    X, Y = np.meshgrid(g["m"], g["n"])
    U0 = Ub.copy()
    _add_pulse(U0, X, Y, 0.0, 0.0, amp_rel * p0, half_width)
    return Case(name=f"euler-pulse-{N}", imax=N, jmax=N, nk=nk, D=D, dm=g["dm"], dn=g["dn"], dt=dt,
                sigmaf=sigmaf, m=g["m"], n=g["n"], meshx=g["meshx"], meshy=g["meshy"],
                sigmax=g["sigmax"], sigmay=g["sigmay"], rhob=rhob, psib=psib, Ub=Ub, U0=U0,
                Re=1e300, inviscid=True)


def pulse_vortex_case(N: int = 4096, D: int = 40, M: float = 0.5, Re_pulse: float = 1.0e4,
                      half_width: float = 4.0, vortex_strength: float = 0.1,
                      vortex_radius: float = 10.0, vortex_offset: float = 200.0,
                      sigmaf: float = 0.2, nk: int = 4, Pr: float = 0.71) -> Case:
    """2-D Navier-Stokes on an N x N uniform grid (dx=1) with uniform mean flow
    M: Gaussian acoustic pulse (1e-3 p0, half-width 4) at the centre plus an
    isentropic vortex (strength 0.1 c0, radius 10) 200 points upstream.
    Re based on the pulse half-width -> Re per unit length = Re_pulse/b."""
    half = (N - 1) / 2.0
    g, Ub, rhob, psib, dt, p0 = _uniform_case("ns", N, half, D, M, sigmaf, nk, Re_pulse / half_width,
                                               Pr, False)
    X, Y = np.meshgrid(g["m"], g["n"])
    U0 = Ub.copy()
    # isentropic vortex: u_theta = eps r/R exp((1-r^2/R^2)/2), T = 1 - (g-1)/2 eps^2 exp(1-r^2/R^2)
    xc, yc = -vortex_offset, 0.0
    R = vortex_radius
    r2 = ((X - xc) ** 2 + (Y - yc) ** 2) / (R * R)
    e = np.exp(0.5 * (1.0 - r2))
    eps = vortex_strength
    U0[..., 1] = M - eps * (Y - yc) / R * e
    U0[..., 2] = eps * (X - xc) / R * e
    T = 1.0 - 0.5 * (GAMMA - 1.0) * eps * eps * e * e
    rho = T ** (1.0 / (GAMMA - 1.0))
    U0[..., 0] = rho
    U0[..., 3] = rho * T / GAMMA
    _add_pulse(U0, X, Y, 0.0, 0.0, 1e-3 * p0, half_width)
    return Case(name=f"ns-pulse-vortex-{N}", imax=N, jmax=N, nk=nk, D=D, dm=g["dm"], dn=g["dn"],
                dt=dt, sigmaf=sigmaf, m=g["m"], n=g["n"], meshx=g["meshx"], meshy=g["meshy"],
                sigmax=g["sigmax"], sigmay=g["sigmay"], rhob=rhob, psib=psib, Ub=Ub, U0=U0,
                Re=Re_pulse / half_width, Pr=Pr)


def batch_sweep_cases(ncases: int = 256, seed: int = 0, sizes=(128, 192, 256, 384, 512), D: int = 20,
                      M_range=(0.0, 0.6), amp_range=(1e-4, 1e-2), hw_range=(2.0, 6.0),
                      max_points: int | None = None) -> list:
    """Seeded sweep of independent 2-D Euler Gaussian-pulse cases: square
    grids drawn uniformly from `sizes`, uniform mean flow M ~ U(M_range),
    amplitude log-uniform in amp_range (relative to p0), half-width
    ~ U(hw_range) grid spacings.  `max_points` (optional) redraws sizes so
    the total point count stays below it."""
    rng = np.random.default_rng(seed)
    out = []
    total = 0
    for c in range(ncases):
        N = int(rng.choice(sizes))
        M = float(rng.uniform(*M_range))
        amp = float(np.exp(rng.uniform(np.log(amp_range[0]), np.log(amp_range[1]))))
        hw = float(rng.uniform(*hw_range))
        if max_points is not None and total + N * N > max_points:
            N = int(min(sizes))
        total += N * N
        case = acoustic_pulse_case(N, D=D, amp_rel=amp, half_width=hw, u0=M)
        case.name = f"sweep-{c}-{N}-M{M:.3f}"
        out.append(case)
    return out


def adjoint_case(N: int = 1024, D: int = 40, M: float = 0.3, ngauss: int = 16, seed: int = 5,
                 amp_sigma: float = 1e-3, hw_range=(3.0, 8.0), inner: float = 0.6) -> Case:
    """2-D Euler on an N x N unit grid with uniform mean flow M; the initial
    pressure perturbation is a seeded sum of `ngauss` Gaussians with
    amplitudes ~ N(0, amp_sigma^2), centres uniform in the inner `inner`
    fraction of the domain and half-widths ~ U(hw_range).  Density and
    velocity start at the base state.  Returns the case with U0 set; the
    perturbation itself is `adjoint_pressure(case_seed...)`."""
    case = acoustic_pulse_case(N, D=D, amp_rel=0.0, u0=M)
    dp = adjoint_pressure(N, ngauss, seed, amp_sigma, hw_range, inner)
    case.U0 = case.Ub.copy()
    case.U0[..., 3] += dp
    case.name = f"euler-adjoint-{N}-M{M}"
    return case


def adjoint_pressure(N: int, ngauss: int = 16, seed: int = 5, amp_sigma: float = 1e-3,
                     hw_range=(3.0, 8.0), inner: float = 0.6) -> np.ndarray:
    """Seeded sum of Gaussians on an N x N unit grid centred on 0, shape (N, N)."""
    rng = np.random.default_rng(seed)
    half = (N - 1) / 2.0
    x = np.arange(N) - half
    X, Y = np.meshgrid(x, x)
    dp = np.zeros((N, N))
    for _ in range(ngauss):
        a = rng.normal(0.0, amp_sigma)
        xc, yc = rng.uniform(-inner * half, inner * half, size=2)
        b = rng.uniform(*hw_range)
        dp += a * np.exp(-math.log(2.0) * ((X - xc) ** 2 + (Y - yc) ** 2) / (b * b))
    return dp


def observer_arc(N: int, radius: float, theta0: float = 0.0, theta1: float = math.pi) -> np.ndarray:
    """Grid points (1-based (i, j), each once) nearest to an arc of the given
    radius about the centre of an N x N unit grid, sampled at one-cell
    spacing from angle theta0 to theta1."""
    half = (N - 1) / 2.0
    nth = max(2, int(math.ceil(radius * (theta1 - theta0))) + 1)
    pts = []
    seen = set()
    for t in np.linspace(theta0, theta1, nth):
        i = int(round(half + radius * math.cos(t))) + 1
        j = int(round(half + radius * math.sin(t))) + 1
        if 1 <= i <= N and 1 <= j <= N and (i, j) not in seen:
            seen.add((i, j))
            pts.append((i, j))
    return np.asarray(pts, dtype=np.int32)


def tanh_stretched_y(jmax: int, Ly: float, ratio: float):
    """Rows y_j = (Ly/2) atanh(eta_j tanh b) / b on the uniform computational
    coordinate eta in [-1, 1]; the spacing grows from the centre to the edges
    by `ratio` (= cosh^2 b).  Returns (eta, y, dn, meshy) with meshy = d eta / dy,
    the metric factor the reference multiplies y-derivatives by."""
    b = math.acosh(math.sqrt(ratio))
    eta = np.linspace(-1.0, 1.0, jmax)
    dn = 2.0 / (jmax - 1)
    tb = math.tanh(b)
    y = 0.5 * Ly * np.arctanh(eta * tb) / b
    dydeta = 0.5 * Ly * tb / (b * (1.0 - (eta * tb) ** 2))
    return eta, y, dn, 1.0 / dydeta


def mixing_layer_case(imax: int = 2048, jmax: int = 1024, D: int = 60, ratio: float = 10.0, delta: float = 1.0,
                      M: float = 0.5, Re_delta: float = 1000.0, Pr: float = 0.72, amp: float = 1e-3,
                      seed: int = 3, cells_per_delta: float = 20.0, nk: int = 4, periodic: bool = True) -> Case:
    """Temporal compressible mixing layer: x periodic and uniform (spacing
    equal to the finest y spacing, delta / cells_per_delta; period imax
    spacings), tanh-stretched y with spacing ratio `ratio`, metric factors
    applied everywhere; base flow u = 0.5 M tanh(2y/delta) with the
    Crocco-Busemann temperature (mixing_layer_fields, init.f90:368-404);
    seeded perturbation v' = amp N(0,1) exp(-(y/delta)^2); Re based on
    delta; D-row sponge buffers at the bottom and the top (the right/top
    profile of init.f90:158-206, mirrored at the bottom).  periodic=False
    keeps the reference's x treatment instead: one-sided x stencils, inflow
    and outflow columns and D-column x buffers (the left one with the
    reference's sigma = 0, init.f90:168)."""
    rng = np.random.default_rng(seed)
    b = math.acosh(math.sqrt(ratio))
    dn = 2.0 / (jmax - 1)
    dy_min = delta / cells_per_delta
    Ly = 2.0 * dy_min * b / (math.tanh(b) * dn)
    eta, y, dn, meshy = tanh_stretched_y(jmax, Ly, ratio)
    dm = dy_min
    half = 0.5 * (imax - 1) * dm
    x = np.linspace(-half, half, imax)
    g = grid_profiles(imax, jmax, -half, half, -1.0, 1.0, D, 0.01, 20.0, sponge_left_bottom=periodic)
    rhob, psib, Ub = mixing_layer_fields(y, imax, nk, 0.5 * M, -0.5 * M, 1.0, 1.0, delta=delta)
    U0 = Ub.copy()
    U0[..., 2] += amp * rng.standard_normal((jmax, imax)) * np.exp(-(y / delta) ** 2)[:, None]
    dt = 0.5 * dy_min / (1.0 + 0.5 * M)
    name = f"mixing-layer-{imax}x{jmax}" + ("-periodic" if periodic else "")
    return Case(name=name, imax=imax, jmax=jmax, nk=nk, D=D, dm=dm, dn=dn, dt=dt,
                sigmaf=0.2, m=x, n=eta, meshx=np.ones(imax), meshy=meshy, sigmax=g["sigmax"], sigmay=g["sigmay"],
                rhob=rhob, psib=psib, Ub=Ub, U0=U0, Re=Re_delta / delta, Pr=Pr, inviscid=False,
                metric_everywhere=True, Tref=273.15, periodic_x=periodic)


def mixing_layer_columns(case: Case):
    """1-based column range the x-averaged statistics use: every column
    when x is periodic, else the columns between the x buffers."""
    return (1, case.imax) if case.periodic_x else (case.D + 1, case.imax - case.D)


def mixing_layer_rows(case: Case, ratio: float = 10.0) -> np.ndarray:
    """Physical y of the rows of a mixing_layer_case (its finest spacing is dm)."""
    b = math.acosh(math.sqrt(ratio))
    Ly = 2.0 * case.dm * b / (math.tanh(b) * case.dn)
    return tanh_stretched_y(case.jmax, Ly, ratio)[1]


def momentum_thickness(ubar: np.ndarray, y: np.ndarray, U1: float, U2: float, rho: np.ndarray | None = None) -> float:
    """theta = int rho (U1 - u)(u - U2) / (rho_ref (U1 - U2)^2) dy over the
    x-averaged profile (trapezoid rule on the physical rows); rho optional."""
    f = (U1 - ubar) * (ubar - U2) / (U1 - U2) ** 2
    if rho is not None:
        f = f * rho
    return float(np.trapezoid(f, y))


def random_case(imax: int, jmax: int, nk: int = 4, D: int = 6, seed: int = 0, Re: float = 500.0,
                amp: float = 1e-2, dt: float = 0.2, sigmaf: float = 0.2, inviscid: bool = False,
                mesh_random: bool = True) -> Case:
    """Small test case with a mixing-layer base flow, a seeded smooth+random
    perturbation, random stretch factors and sponge profiles on all sides."""
    rng = np.random.default_rng(seed)
    g = grid_profiles(imax, jmax, -0.5 * (imax - 1), 0.5 * (imax - 1), -0.5 * (jmax - 1),
                      0.5 * (jmax - 1), D, 0.01, 20.0, sponge_left_bottom=True)
    rhob, psib, Ub = mixing_layer_fields(g["n"] / (0.25 * jmax), imax, nk, 0.6, 0.1, 1.0, 0.8)
    if mesh_random:
        g["meshx"] = 1.0 + 0.3 * rng.random(imax)
        g["meshy"] = 1.0 + 0.3 * rng.random(jmax)
    U0 = Ub.copy()
    X, Y = np.meshgrid(np.arange(imax), np.arange(jmax))
    for k in range(nk):
        smooth = np.sin(2 * np.pi * (X * (k + 1) / imax + Y * (k + 2) / jmax))
        U0[..., k] += amp * (smooth + 0.3 * rng.standard_normal((jmax, imax)))
    U0[..., 0] = np.abs(U0[..., 0]) + 0.5   # keep rho > 0
    U0[..., 3] = np.abs(U0[..., 3]) + 0.3   # keep p > 0
    return Case(name=f"random-{imax}x{jmax}-nk{nk}-s{seed}", imax=imax, jmax=jmax, nk=nk, D=D,
                dm=g["dm"], dn=g["dn"], dt=dt, sigmaf=sigmaf, m=g["m"], n=g["n"], meshx=g["meshx"],
                meshy=g["meshy"], sigmax=g["sigmax"], sigmay=g["sigmay"], rhob=rhob, psib=psib,
                Ub=Ub, U0=U0, Re=Re, inviscid=inviscid, Tref=273.15)
