"""Initial conditions and descriptors for the benchmark and verification problems.

Initial fields are built once on the host in FP64 and the same bits are
uploaded to every implementation.  Transcendental factors are separable in
these problems, so they are evaluated per axis with the C library (``math``)
-- the same libm the reference's generated initialisation kernel calls -- and
combined with IEEE multiplications in the order the reference's generated
initialisation kernel uses (apps/performance_benchmark/TENO6/
opensbliblock00_kernels.h:3-46).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import native as nat


def _axis(fn, xs):
    return np.array([fn(float(x)) for x in xs], dtype=np.float64)


@dataclass
class Problem:
    name: str
    desc: nat.ProblemDesc
    q0: np.ndarray
    steps: int

    @property
    def work_units(self) -> float:
        """Grid-point RK-stage updates of the full run."""
        return float(self.desc.npoints) * 3.0 * self.steps


# ---------------------------------------------------------------------------
# 1-D Sod shock tube (apps/Sod_shock_tube/Sod_shock_tube.py:72-81)
# ---------------------------------------------------------------------------
def sod(n: int = 1024, steps: int = 1000, dt: float = 2.0e-4, precision: int = nat.PREC_FP64,
        boundaries: str = "zero_gradient") -> Problem:
    """``boundaries``: 'zero_gradient' (apps/Sod_shock_tube), 'dirichlet' (the
    left and right states held at both ends,
    tests/verification_apps/sod_shock_tube/verify_sod_shock_tube.py:96-109),
    'extrapolation' (order 0) or 'linear' (order-1 extrapolation)."""
    gama = 1.4
    delta = 1.0 / (n - 1)
    x = delta * np.arange(n, dtype=np.float64)
    left = x < 0.5
    d = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.10)
    u0 = np.zeros(n)
    q = np.empty((3, 1, 1, n))
    q[0, 0, 0] = d
    q[1, 0, 0] = d * u0
    q[2, 0, 0] = p / (gama - 1.0) + 0.5 * d * (u0 ** 2.0)
    desc = nat.ProblemDesc(
        ndim=1, n=[n, 1, 1], delta=[delta, 1.0, 1.0], conv_scheme=nat.CONV_WENO_Z, order=5,
        averaging=nat.AVG_SIMPLE, time_scheme=nat.TIME_RKLS3, precision=precision,
        bc=[[nat.BC_ZERO_GRADIENT, nat.BC_ZERO_GRADIENT]], gama=gama, dt=dt)
    if boundaries == "dirichlet":
        desc.bc = [[nat.BC_DIRICHLET, nat.BC_DIRICHLET]]
        desc.bc_value = {(0, 0): [1.0, 1.0 * 0.0, 1.0 / (gama - 1.0) + 0.5 * 1.0 * (0.0 ** 2.0)],
                         (0, 1): [0.125, 0.125 * 0.0, 0.1 / (gama - 1.0) + 0.5 * 0.125 * (0.0 ** 2.0)]}
    elif boundaries == "extrapolation":
        desc.bc = [[nat.BC_EXTRAPOLATION0, nat.BC_EXTRAPOLATION0]]
    elif boundaries == "linear":
        desc.bc = [[nat.BC_EXTRAPOLATION1, nat.BC_EXTRAPOLATION1]]
    elif boundaries != "zero_gradient":
        raise ValueError(f"unknown Sod boundaries {boundaries!r}")
    return Problem("sod", desc, q, steps)


def sod_filtered(n: int = 200, steps: int = 1000, dt: float = 2.0e-4) -> Problem:
    """The Sod verification app's default method
    (tests/verification_apps/sod_shock_tube/verify_sod_shock_tube.py:37-191):
    Central(4) Feiereisen split (StoreSome 'u0'), inviscid, the WENO-Z filter
    after each step, RungeKuttaLS(3, 'SSP'), Dirichlet faces holding the left
    and right states (so the central derivatives next to them use Carpenter's
    closure), N = 200, dt = 2e-4, 1000 steps."""
    pr = sod(n=n, steps=steps, dt=dt, boundaries="dirichlet")
    # RungeKuttaLS(3, formulation='SSP'): the 3-stage 2N-storage SSP coefficients
    from .surface import RungeKuttaLS
    ssp = RungeKuttaLS(3, formulation="SSP")
    d = pr.desc
    d.conv_scheme, d.order = nat.CONV_CENTRAL_FEIEREISEN, 4
    d.rk_a, d.rk_b = list(ssp.rk_a), list(ssp.rk_b)
    d.shock_filter = nat.FILTER_WENOZ
    return Problem(f"sod_filtered_{n}", d, pr.q0, steps)


# ---------------------------------------------------------------------------
# 3-D Taylor-Green vortex (apps/taylor_green_vortex/taylor_green_vortex.py,
# apps/compressible_taylor_green_vortex/compressible_TGV.py)
# ---------------------------------------------------------------------------
def _tgv_fields(n: Tuple[int, int, int], delta: Tuple[float, float, float], gama: float, minf: float,
                offset: Tuple[int, int, int] = (0, 0, 0)):
    xs = [delta[d] * (np.arange(n[d]) + offset[d]) for d in range(3)]
    s = [_axis(math.sin, x) for x in xs]
    c = [_axis(math.cos, x) for x in xs]
    c2 = [_axis(lambda v: math.cos(2.0 * v), x) for x in xs]
    S0, C0, C20 = s[0][None, None, :], c[0][None, None, :], c2[0][None, None, :]
    S1, C1, C21 = s[1][None, :, None], c[1][None, :, None], c2[1][None, :, None]
    C2, C22 = c[2][:, None, None], c2[2][:, None, None]
    # u0 = cos(x1)*cos(x2)*sin(x0); u1 = -cos(x0)*cos(x2)*sin(x1); u2 = 0
    u0 = (C1 * C2) * S0
    u1 = -((C0 * C2) * S1)
    u2 = np.zeros_like(u0)
    # p = (2.0 + cos(2.0*x2))*(0.0625*cos(2.0*x0) + 0.0625*cos(2.0*x1)) + 1.0/((Minf*Minf)*gama)
    p = (2.0 + C22) * (0.0625 * C20 + 0.0625 * C21) + 1.0 / ((minf * minf) * gama)
    r = gama * p * (minf * minf)
    rhoE = p / (-1 + gama) + 0.5 * r * ((u0 * u0) + (u1 * u1) + (u2 * u2))
    q = np.empty((5, n[2], n[1], n[0]))
    q[0] = r
    q[1] = r * u0
    q[2] = r * u1
    q[3] = r * u2
    q[4] = rhoE
    return q


# RK3 with 4th-order central differences is stable for |lambda| dt <= sqrt(3); at
# Ma = 0.1 the acoustic eigenvalue a*sqrt(3)*1.372/Delta makes dt = 5e-3 at 128^3
# diverge (by step ~60), so the central runs default to the reference app's own
# dt = 0.003385 at 128^3 (apps/taylor_green_vortex/taylor_green_vortex.py:12),
# scaled with the grid spacing.
TGV_CENTRAL_DT_128 = 0.003385


def tgv(n: int = 128, steps: int = 200, dt: Optional[float] = None, minf: float = 0.1, re: float = 1600.0,
        pr: float = 0.71, scheme: str = "central", precision: int = nat.PREC_FP64,
        teno_order: int = 5, teno_ct: float = 1.0e-5, teno_eps: float = 1.0e-40, cfl: float = 0.0,
        averaging: int = nat.AVG_ROE) -> Problem:
    """Periodic [0, 2pi)^3 Taylor-Green vortex.

    scheme='central': NS_Split('Feiereisen', viscosity='constant') + StoreSome(4, 'u0 u1 u2')
    scheme='kgp'    : NS_Split('KGP', viscosity='constant', energy_formulation='enthalpy') + StoreSome(4)
    scheme='teno'   : LFTeno(order, averaging=RoeAverage) + central Sutherland viscous terms
    """
    gama = 1.4
    delta = 2.0 * math.pi / n
    if dt is None:
        dt = TGV_CENTRAL_DT_128 * 128.0 / n
    q = _tgv_fields((n, n, n), (delta, delta, delta), gama, minf)
    if scheme == "central":
        conv, order, visc = nat.CONV_CENTRAL_FEIEREISEN, 4, nat.VISC_CONSTANT
    elif scheme == "kgp":
        conv, order, visc = nat.CONV_CENTRAL_KGP, 4, nat.VISC_CONSTANT
    elif scheme == "teno":
        conv, order, visc = nat.CONV_TENO, teno_order, nat.VISC_SUTHERLAND
    elif scheme == "weno":
        conv, order, visc = nat.CONV_WENO_Z, 5, nat.VISC_SUTHERLAND
    else:
        raise ValueError(scheme)
    desc = nat.ProblemDesc(
        ndim=3, n=[n, n, n], delta=[delta] * 3, conv_scheme=conv, order=order, averaging=averaging,
        viscous=visc, time_scheme=nat.TIME_RKLS3, precision=precision,
        bc=[[nat.BC_PERIODIC] * 2] * 3, gama=gama, minf=minf, re=re, pr=pr,
        teno_eps=teno_eps, teno_ct=teno_ct, dt=dt, cfl=cfl)
    return Problem(f"tgv_{scheme}_{n}", desc, q, steps)


def tgv_central4_weno5(n: int = 128, steps: int = 500, dt: Optional[float] = None) -> Problem:
    """The reference's other performance-benchmark method
    (apps/performance_benchmark/Central4-WENO5/opensbli.cpp:36-110,
    apps/performance_benchmark/README.MD:27-29): KGP cubic split (enthalpy
    form), Sutherland viscous terms, RK-LS3, and after each step the WENO3-Z
    filter with RoeAverage, LLF and the Ducros hybrid selection
    (Ducros_check = Ducros_select = 0.05).  dt = 2 * 0.00084625 at the
    validation size 128^3 ("va"), 0.00084625 * 256 / n otherwise ("ss")."""
    if dt is None:
        dt = 0.00084625 * 2.0 if n == 128 else 0.00084625 * 256.0 / n
    pr = tgv(n=n, steps=steps, dt=dt, minf=1.25, re=1600.0, pr=0.71, scheme="kgp")
    d = pr.desc
    d.viscous = nat.VISC_SUTHERLAND
    d.shock_filter, d.filter_order, d.filter_sensor = nat.FILTER_WENOZ, 3, nat.SENSOR_DUCROS
    d.ducros_check = d.ducros_select = 0.05
    return Problem(f"tgv_central4_weno5_{n}", d, pr.q0, steps)


def tgv_teno6_validation(n: int = 128, steps: int = 500) -> Problem:
    """The reference's performance-benchmark validation run ("va"):
    apps/performance_benchmark/TENO6/opensbli.cpp:36-48, 94-140 -- TENO6,
    RoeAverage, LLF, Sutherland viscosity, RK-LS3, dt = 2*0.00084625."""
    return tgv(n=n, steps=steps, dt=0.00084625 * 2.0, minf=1.25, re=1600.0, pr=0.71, scheme="teno",
               teno_order=6, teno_ct=1e-6, teno_eps=1e-40)


def tg_sym(n: int = 33, steps: int = 1) -> Problem:
    """Triply-symmetric Taylor-Green verification app
    (tests/verification_apps/tg_sym/verify_tg_sym.py:12-178): Central(4) with
    Skew convective terms, SymmetryBC on all faces, classic RK3, [0, pi]^3."""
    gama, minf, re, pr = 1.4, 0.1, 800.0, 0.71
    delta = math.pi / (n - 1)
    q = _tgv_fields((n, n, n), (delta, delta, delta), gama, minf)
    desc = nat.ProblemDesc(
        ndim=3, n=[n, n, n], delta=[delta] * 3, conv_scheme=nat.CONV_CENTRAL_SKEW, order=4,
        viscous=nat.VISC_CONSTANT, time_scheme=nat.TIME_RK3, precision=nat.PREC_FP64,
        bc=[[nat.BC_SYMMETRY] * 2] * 3, gama=gama, minf=minf, re=re, pr=pr, dt=0.005)
    return Problem("tg_sym", desc, q, steps)


# ---------------------------------------------------------------------------
# Periodic entropy wave under the WENO filter (opensbli/filters/WENO_filter.py:406)
# ---------------------------------------------------------------------------
def entropy_wave(n: int = 64, ndim: int = 2, profile: str = "step", steps: int = 20, filtered: bool = True,
                 scheme: str = "central", velocity: float = 1.0, cfl: float = 0.0) -> Problem:
    """A density profile advected at constant velocity and pressure on the
    periodic unit cube: 'step' (a top hat, density 1 / 0.5 / 1 over x <
    0.25, 0.25 <= x < 0.75, x >= 0.75) or 'sine' (1 + 0.2 sin 2 pi x).  The
    profile varies along x; in 2-D and 3-D the other directions advect at
    the same speed with the profile constant along them.  ``scheme``:
    'central' (Central(4) Feiereisen, inviscid -- the base scheme the filter
    is designed for), 'kgp' (the enthalpy split form) or 'weno' (WENO-Z,
    simple averaging).  ``filtered``: the WENO-Z filter after each step."""
    gama, pr0 = 1.4, 1.0
    delta = 1.0 / n
    x = delta * np.arange(n, dtype=np.float64)
    if profile == "step":
        d1 = np.where((x >= 0.25) & (x < 0.75), 0.5, 1.0)
    elif profile == "sine":
        d1 = 1.0 + 0.2 * _axis(lambda v: math.sin(2.0 * math.pi * v), x)
    else:
        raise ValueError(f"unknown profile {profile!r}")
    shape = [1, 1, 1]
    for d in range(ndim):
        shape[2 - d] = n
    rho = np.broadcast_to(d1[None, None, :], tuple(shape)).copy()
    nv = ndim + 2
    q = np.empty((nv,) + tuple(shape))
    q[0] = rho
    for d in range(ndim):
        q[1 + d] = rho * velocity
    q[nv - 1] = pr0 / (gama - 1.0) + 0.5 * rho * (ndim * velocity * velocity)
    conv = {"central": nat.CONV_CENTRAL_FEIEREISEN, "kgp": nat.CONV_CENTRAL_KGP, "weno": nat.CONV_WENO_Z}[scheme]
    desc = nat.ProblemDesc(
        ndim=ndim, n=[n if d < ndim else 1 for d in range(3)], delta=[delta if d < ndim else 1.0 for d in range(3)],
        conv_scheme=conv, order=5 if scheme == "weno" else 4, averaging=nat.AVG_SIMPLE,
        viscous=nat.VISC_NONE, time_scheme=nat.TIME_RKLS3, precision=nat.PREC_FP64,
        bc=[[nat.BC_PERIODIC, nat.BC_PERIODIC]] * ndim, gama=gama, dt=0.0 if cfl > 0 else 0.2 * delta, cfl=cfl,
        shock_filter=nat.FILTER_WENOZ if filtered else nat.FILTER_NONE)
    return Problem(f"entropy_{profile}_{ndim}d_{n}_{scheme}{'_filtered' if filtered else ''}", desc, q, steps)


# ---------------------------------------------------------------------------
# 2-D Kelvin-Helmholtz instance (config: periodic x, slip walls y)
# ---------------------------------------------------------------------------
def kelvin_helmholtz(n: int = 128, amplitude: float = 0.02, mode: int = 2, steps: int = 500,
                     precision: int = nat.PREC_FP64, dt: Optional[float] = None) -> Problem:
    gama = 1.4
    delta = 1.0 / n
    x = -0.5 + delta * np.arange(n)
    y = -0.5 + delta * np.arange(n)
    dl, sg = 0.025, 0.05
    th = _axis(lambda v: math.tanh((v + 0.25) / dl), y) - _axis(lambda v: math.tanh((v - 0.25) / dl), y)
    rho_y = 1.0 + 0.5 * th
    u_y = -0.5 + 0.5 * th
    g_y = _axis(lambda v: math.exp(-((v - 0.25) / sg) ** 2), y) + _axis(lambda v: math.exp(-((v + 0.25) / sg) ** 2), y)
    s_x = _axis(lambda v: math.sin(2.0 * math.pi * mode * (v + 0.5)), x)
    rho = np.broadcast_to(rho_y[:, None], (n, n)).copy()
    u = np.broadcast_to(u_y[:, None], (n, n)).copy()
    v = amplitude * s_x[None, :] * g_y[:, None]
    p = 2.5
    q = np.empty((4, 1, n, n))
    q[0, 0] = rho
    q[1, 0] = rho * u
    q[2, 0] = rho * v
    q[3, 0] = p / (gama - 1.0) + 0.5 * rho * (u * u + v * v)
    desc = nat.ProblemDesc(
        ndim=2, n=[n, n, 1], delta=[delta, delta, 1.0], conv_scheme=nat.CONV_WENO_Z, order=5,
        averaging=nat.AVG_ROE, time_scheme=nat.TIME_RKLS3, precision=precision,
        bc=[[nat.BC_PERIODIC, nat.BC_PERIODIC], [nat.BC_INVISCID_WALL, nat.BC_INVISCID_WALL]],
        gama=gama, dt=dt if dt is not None else 0.2 / n)
    return Problem(f"kh_{n}", desc, q, steps)


def kh_ensemble(count: int = 64, seed: int = 20261203, sizes=(128, 256, 512, 1024), steps: int = 500,
                precision: int = nat.PREC_MIXED):
    """Seeded ensemble of Kelvin-Helmholtz instances of uneven sizes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    probs = []
    for i in range(count):
        n = int(rng.choice(np.array(sizes)))
        amp = float(rng.uniform(0.01, 0.05))
        mode = int(rng.integers(1, 5))
        pr = kelvin_helmholtz(n=n, amplitude=amp, mode=mode, steps=steps, precision=precision)
        pr.name = f"kh{i}_{n}"
        probs.append(pr)
    return probs


def tgv_ensemble(count: int = 48, seed: int = 20261204, sizes=(32, 48, 64, 96), steps: int = 200):
    """Seeded ensemble of 3-D Taylor-Green vortices of uneven sizes (FP64,
    Feiereisen split + constant viscosity, every face periodic): a Mach number in
    [0.1, 0.3] and a Reynolds number in [400, 1600] per instance."""
    rng = np.random.Generator(np.random.PCG64(seed))
    probs = []
    for i in range(count):
        n = int(rng.choice(np.array(sizes)))
        minf = float(rng.uniform(0.1, 0.3))
        re = float(rng.uniform(400.0, 1600.0))
        pr = tgv(n=n, steps=steps, minf=minf, re=re)
        pr.name = f"tgv{i}_{n}"
        probs.append(pr)
    return probs


# ---------------------------------------------------------------------------
# 2-D four-quadrant Riemann problem (serving stream)
# ---------------------------------------------------------------------------
def riemann_2d(n: int = 256, states=None, steps: int = 100, precision: int = nat.PREC_FP64,
               dt: Optional[float] = None, boundaries: str = "zero_gradient") -> Problem:
    """``boundaries``: 'zero_gradient' on all four faces, or 'mixed': x0 = 0
    Dirichlet at the lower-left state, x0 = 1 linear extrapolation, x1 = 0
    order-0 extrapolation, x1 = 1 symmetry (every face type that reads or writes
    the corner halos of the other direction)."""
    gama = 1.4
    if states is None:
        # (rho, u, v, p) for quadrants (x>.5,y>.5), (x<.5,y>.5), (x<.5,y<.5), (x>.5,y<.5)
        states = [(1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3), (0.138, 1.206, 1.206, 0.029),
                  (0.5323, 0.0, 1.206, 0.3)]
    delta = 1.0 / (n - 1)
    x = delta * np.arange(n)
    X = x[None, :]
    Y = x[:, None]
    q = np.empty((4, 1, n, n))
    rho = np.empty((n, n)); u = np.empty((n, n)); v = np.empty((n, n)); p = np.empty((n, n))
    masks = [(X >= 0.5) & (Y >= 0.5), (X < 0.5) & (Y >= 0.5), (X < 0.5) & (Y < 0.5), (X >= 0.5) & (Y < 0.5)]
    for m, (r_, u_, v_, p_) in zip(masks, states):
        mm = np.broadcast_to(m, (n, n))
        rho[mm] = r_; u[mm] = u_; v[mm] = v_; p[mm] = p_
    q[0, 0] = rho
    q[1, 0] = rho * u
    q[2, 0] = rho * v
    q[3, 0] = p / (gama - 1.0) + 0.5 * rho * (u * u + v * v)
    if dt is None:
        smax = max(abs(s[1]) + abs(s[2]) + 2 * math.sqrt(gama * s[3] / s[0]) for s in states)
        dt = 0.25 * delta / smax
    desc = nat.ProblemDesc(
        ndim=2, n=[n, n, 1], delta=[delta, delta, 1.0], conv_scheme=nat.CONV_WENO_Z, order=5,
        averaging=nat.AVG_ROE, time_scheme=nat.TIME_RKLS3, precision=precision,
        bc=[[nat.BC_ZERO_GRADIENT] * 2] * 2, gama=gama, dt=dt)
    if boundaries == "mixed":
        desc.bc = [[nat.BC_DIRICHLET, nat.BC_EXTRAPOLATION1], [nat.BC_EXTRAPOLATION0, nat.BC_SYMMETRY]]
        r_, u_, v_, p_ = states[2]
        desc.bc_value = {(0, 0): [r_, r_ * u_, r_ * v_, p_ / (gama - 1.0) + 0.5 * r_ * (u_ * u_ + v_ * v_)]}
        return Problem(f"riemann_mixed_{n}", desc, q, steps)
    if boundaries != "zero_gradient":
        raise ValueError(f"unknown boundaries {boundaries!r}")
    return Problem(f"riemann_{n}", desc, q, steps)
