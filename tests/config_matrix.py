"""Seeded random configurations across the hot path's options, for the
parity sweep (tests/test_gpu_matrix.py) and its CPU check.

Each case is a small grid (8..20 points per direction) with a smooth random
state (positive density and pressure, velocities below the sound speed), a
scheme, boundary faces, time integrator, precision-free options and filter
drawn from what both libraries accept (including HLLC / HLLC-LM fluxes with
WENO-Z)."""
import math

import numpy as np

from opensbli_amd import native as nat, problems as P

_SHOCK = [(nat.CONV_WENO_Z, 5), (nat.CONV_TENO, 5), (nat.CONV_TENO, 6)]
_CENTRAL = [nat.CONV_CENTRAL_FEIEREISEN, nat.CONV_CENTRAL_SKEW, nat.CONV_CENTRAL_KGP]
_NONPERIODIC = [nat.BC_ZERO_GRADIENT, nat.BC_INVISCID_WALL, nat.BC_SYMMETRY, nat.BC_EXTRAPOLATION0,
                nat.BC_EXTRAPOLATION1, nat.BC_DIRICHLET]


def _state(rng, ndim, n, gama):
    shape = [1, 1, 1]
    for d in range(ndim):
        shape[2 - d] = n[d]
    axes = [np.arange(n[d]) * (2 * math.pi / n[d]) for d in range(ndim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1] if ndim > 1 else [axes[0]]

    def field(amp):
        f = np.zeros(tuple(n[d] for d in range(ndim))[::-1])
        for d in range(ndim):
            k = rng.integers(1, 3)
            f = f + amp * rng.uniform(-1, 1) * np.sin(k * grids[d] + rng.uniform(0, 2 * math.pi))
        return f.reshape(shape)

    rho = 1.0 + field(0.15)
    p = 1.0 + field(0.15)
    a = math.sqrt(gama)
    u = [field(0.25 * a) for _ in range(ndim)]
    q = np.empty((ndim + 2,) + tuple(shape))
    q[0] = rho
    for d in range(ndim):
        q[1 + d] = rho * u[d]
    q[-1] = p / (gama - 1.0) + 0.5 * rho * sum(ud * ud for ud in u)
    return q


def make_case(seed):
    """One valid configuration and its initial state, from `seed`."""
    rng = np.random.default_rng(seed)
    ndim = int(rng.integers(1, 4))
    n = [int(rng.integers(8, 21 if ndim < 3 else 15)) for _ in range(ndim)]
    gama = 1.4
    central = ndim >= 2 and rng.random() < 0.4
    if ndim == 1 and rng.random() < 0.25:
        central = True
    viscous = nat.VISC_NONE
    if ndim >= 2 and rng.random() < 0.5:
        viscous = nat.VISC_CONSTANT if rng.random() < 0.5 else nat.VISC_SUTHERLAND
    if central:
        conv, order = int(rng.choice(_CENTRAL)), 4
        if ndim == 1:
            conv = int(rng.choice([nat.CONV_CENTRAL_FEIEREISEN, nat.CONV_CENTRAL_SKEW]))
    else:
        conv, order = _SHOCK[int(rng.integers(0, 3))]
    bc = []
    for d in range(ndim):
        if rng.random() < 0.5:
            bc.append([nat.BC_PERIODIC, nat.BC_PERIODIC])
        else:
            bc.append([int(rng.choice(_NONPERIODIC)), int(rng.choice(_NONPERIODIC))])
    q = _state(rng, ndim, n, gama)
    bc_value = {}
    for d in range(ndim):
        for s in range(2):
            if bc[d][s] == nat.BC_DIRICHLET:
                bc_value[(d, s)] = [float(q[v].mean()) for v in range(ndim + 2)]
    h = [2 * math.pi / n[d] for d in range(ndim)]
    desc = nat.ProblemDesc(ndim=ndim, n=n + [1] * (3 - ndim), delta=h + [1.0] * (3 - ndim), conv_scheme=conv, order=order,
                           averaging=int(rng.choice([nat.AVG_SIMPLE, nat.AVG_ROE])), viscous=viscous,
                           bc=bc, bc_value=bc_value, gama=gama, minf=0.5, re=200.0, pr=0.71)
    ts = rng.random()
    if ts < 0.25:
        desc.time_scheme = nat.TIME_RK3
    elif ts < 0.5:
        from opensbli_amd.surface import RungeKuttaLS
        rk = RungeKuttaLS(3, stages=int(rng.integers(3, 6)), formulation="SSP")
        desc.rk_a, desc.rk_b = list(rk.rk_a), list(rk.rk_b)
    if rng.random() < 0.4:
        desc.cfl = 0.3
    else:
        desc.dt = 0.2 * min(h) / (1.0 + math.sqrt(gama))
    if not central and rng.random() < 0.3:
        desc.flux_type = nat.FLUX_GLF
    elif conv == nat.CONV_WENO_Z and rng.random() < 0.35:
        desc.flux_type = nat.FLUX_HLLC if rng.random() < 0.5 else nat.FLUX_HLLC_LM
    if desc.flux_type == nat.FLUX_LLF and rng.random() < 0.35:
        desc.shock_filter = nat.FILTER_WENOZ
        desc.filter_order = int(rng.choice([3, 5]))
        if ndim >= 2 and rng.random() < 0.5:
            desc.filter_sensor = nat.SENSOR_DUCROS
    return P.Problem(f"matrix_{seed}", desc, q.reshape(desc.shape), 3)


CASES = list(range(20261201, 20261201 + 72))
