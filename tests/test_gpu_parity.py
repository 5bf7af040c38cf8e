"""HIP library against the CPU oracle on the same inputs (FP64 bar: relative
L2 and L-inf <= 1e-11 per conservative variable, as stated for the FP64 path).

Each case prints its name before touching the GPU so a failing run names it."""
import math
import os

import numpy as np
import pytest

from opensbli_amd import native as nat, problems as P
from helpers import assert_exact_unless_fused, fp64_bar, fused_3d, kept_fastest, rel_errors, run

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)


def _compare(hip, oracle, pr, steps=None, tol=TOL):
    _announce(f"{pr.name} steps={steps if steps is not None else pr.steps} n={list(pr.desc.n)}")
    a, sa = run(hip, pr, steps)
    b, sb = run(oracle, pr, steps)
    l2, li = rel_errors(a, b)
    print(f"   rel L2 {l2.max():.3e}  rel Linf {li.max():.3e}", flush=True)
    assert np.all(np.isfinite(a))
    bound = fp64_bar(a, b, pr.desc, oracle, pr, steps, tol)
    assert sa.iter == sb.iter
    assert math.isclose(sa.time, sb.time, rel_tol=max(1e-12, bound), abs_tol=1e-15)
    return a, b


def _residual_compare(hip, oracle, pr, state=None, tol=TOL):
    _announce(f"residual {pr.name} n={list(pr.desc.n)}")
    q = pr.q0 if state is None else state
    with hip.context(pr.desc) as c:
        c.set_state(q)
        ra = c.residual()
    with oracle.context(pr.desc) as c:
        c.set_state(q)
        rb = c.residual()
    scale = np.array([np.max(np.abs(rb[v])) for v in range(rb.shape[0])])
    err = np.array([np.max(np.abs(ra[v] - rb[v])) for v in range(rb.shape[0])]) / np.maximum(scale, 1e-300)
    print(f"   residual rel Linf {err.max():.3e}", flush=True)
    assert err.max() <= 1e-12, err
    return ra, rb


# --- single right-hand-side evaluations -------------------------------------------
def test_residual_sod(hip, oracle):
    _residual_compare(hip, oracle, P.sod(n=256))


def test_residual_kh(hip, oracle):
    _residual_compare(hip, oracle, P.kelvin_helmholtz(n=64))


def test_residual_riemann(hip, oracle):
    _residual_compare(hip, oracle, P.riemann_2d(n=64))


def test_residual_tgv_central(hip, oracle):
    _residual_compare(hip, oracle, P.tgv(n=24))


def test_residual_tgv_teno5(hip, oracle):
    _residual_compare(hip, oracle, P.tgv(n=24, scheme="teno", minf=1.25))


def test_residual_tgv_teno6(hip, oracle):
    _residual_compare(hip, oracle, P.tgv_teno6_validation(n=24))


def test_residual_tgv_kgp(hip, oracle):
    _residual_compare(hip, oracle, P.tgv(n=24, scheme="kgp"))


def test_kgp_runs(hip, oracle):
    """KGP (enthalpy) with constant-viscosity terms: the separate central kernels in 3-D and 2-D."""
    pr = P.tgv(n=24, steps=5, scheme="kgp")
    a, b = _compare(hip, oracle, pr)
    assert_exact_unless_fused(a, b, pr.desc)
    # a smooth periodic 2-D field with density and pressure variation
    n, gama = 40, 1.4
    h = 2 * math.pi / n
    X, Y = np.meshgrid(h * np.arange(n), h * np.arange(n), indexing="xy")
    rho = 1.0 + 0.2 * np.sin(X) * np.cos(2 * Y)
    u, v = 0.3 * np.sin(X + 0.5) * np.cos(Y), -0.2 * np.cos(X) * np.sin(Y - 0.3)
    p = 1.0 + 0.1 * np.cos(X - Y)
    q = np.stack([rho, rho * u, rho * v, p / (gama - 1.0) + 0.5 * rho * (u * u + v * v)])[:, None]
    desc = nat.ProblemDesc(ndim=2, n=[n, n], delta=[h, h], conv_scheme=nat.CONV_CENTRAL_KGP, order=4,
                           viscous=nat.VISC_CONSTANT, re=100.0, minf=0.3, bc=[[nat.BC_PERIODIC] * 2] * 2,
                           gama=gama, dt=2e-3)
    pr = P.Problem("kgp_2d", desc, q, 20)
    a, b = _compare(hip, oracle, pr)
    assert_exact_unless_fused(a, b, pr.desc)


def test_residual_tg_sym(hip, oracle):
    _residual_compare(hip, oracle, P.tg_sym())


# --- full runs --------------------------------------------------------------------
def test_sod_full_config(hip, oracle):
    """Sod, N=1024, WENO-Z LLF, RK3-LS, dt=2e-4, 1000 steps -- the full configuration."""
    _compare(hip, oracle, P.sod())


def test_tgv_central_small(hip, oracle):
    _compare(hip, oracle, P.tgv(n=32, steps=10))


def test_tgv_teno5_cfl_small(hip, oracle):
    _compare(hip, oracle, P.tgv(n=24, steps=5, scheme="teno", minf=1.25, cfl=0.5, dt=0.0))


def test_tgv_teno6_small(hip, oracle):
    _compare(hip, oracle, P.tgv_teno6_validation(n=24, steps=5))


def test_kh_small(hip, oracle):
    _compare(hip, oracle, P.kelvin_helmholtz(n=64, steps=50))


def test_riemann_small(hip, oracle):
    _compare(hip, oracle, P.riemann_2d(n=64, steps=50))


# --- extrapolation and Dirichlet faces (extrapolation.py:29-58, dirichlet.py:28-41) ---
@pytest.mark.parametrize("boundaries", ["dirichlet", "extrapolation", "linear"])
def test_sod_boundary_variants(hip, oracle, boundaries):
    pr = P.sod(boundaries=boundaries)
    pr.name += "_" + boundaries
    a, b = _compare(hip, oracle, pr)
    assert np.array_equal(a, b)


def test_riemann_mixed_boundaries(hip, oracle):
    _residual_compare(hip, oracle, P.riemann_2d(n=48, boundaries="mixed"))
    a, b = _compare(hip, oracle, P.riemann_2d(n=64, steps=60, boundaries="mixed"))
    assert np.array_equal(a, b)


# --- GLF flux splitting (shock_capturing.py:416-438, :838-857) --------------------
def _glf(pr):
    pr.desc.flux_type = nat.FLUX_GLF
    pr.name += "_glf"
    return pr


def test_glf_sod_teno_dirichlet(hip, oracle):
    """The shock-capturing branch of verify_sod_shock_tube.py: LFTeno(5, GLF), Dirichlet ends."""
    pr = _glf(P.sod(boundaries="dirichlet"))
    pr.desc.conv_scheme = nat.CONV_TENO
    a, b = _compare(hip, oracle, pr)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("maker", [
    lambda: P.kelvin_helmholtz(n=64, steps=40),
    lambda: P.riemann_2d(n=48, steps=40, boundaries="mixed"),
    lambda: P.tgv(n=16, steps=4, scheme="teno", minf=1.25),            # fused face fluxes + viscous
])
def test_glf_parity(hip, oracle, maker):
    pr = _glf(maker())
    _residual_compare(hip, oracle, pr)
    a, b = _compare(hip, oracle, pr)
    if not fused_3d(pr.desc):
        assert np.array_equal(a, b)


def test_3d_shock_extrapolation_dirichlet(hip, oracle):
    """3-D inviscid WENO-Z: extrapolation in x0, Dirichlet/linear in x1, periodic
    x2 -- the BC kernels over both transverse halo ranges."""
    pr = P.tgv(n=16, steps=8, scheme="weno", minf=1.25)
    pr.desc.viscous = nat.VISC_NONE
    pr.desc.bc = [[nat.BC_EXTRAPOLATION0, nat.BC_EXTRAPOLATION1], [nat.BC_DIRICHLET, nat.BC_EXTRAPOLATION1],
                  [nat.BC_PERIODIC, nat.BC_PERIODIC]]
    q0 = pr.q0[:, 0, 0, 0]
    pr.desc.bc_value = {(1, 0): list(q0)}
    pr.name += "_extrap_dirichlet"
    _residual_compare(hip, oracle, pr)
    a, b = _compare(hip, oracle, pr)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("order,stages,formulation", [(3, 5, "SSP"), (4, None, None)])
def test_rkls_variants(hip, oracle, order, stages, formulation):
    """RungeKuttaLS SSP(5,3) and the 4th-order 5-stage scheme through the
    generic 2N-storage path (KH, and the fused 3-D central stage kernel)."""
    from opensbli_amd.surface import RungeKuttaLS
    rk = RungeKuttaLS(order, stages=stages, formulation=formulation)
    for pr in (P.kelvin_helmholtz(n=64, steps=20), P.tgv(n=24, steps=4)):
        pr.desc.rk_a, pr.desc.rk_b = rk.rk_a, rk.rk_b
        pr.name += f"_rk{order}_{len(rk.rk_a)}"
        _compare(hip, oracle, pr)


def test_hip_target_default_library(oracle):
    """HIPTarget drives libosb_hip.so by default (the drop-in for OPSC(alg))."""
    from opensbli_amd.code_generation.hip import HIPTarget
    from opensbli_amd.surface import (LFWeno, RungeKuttaLS, SimpleAverage, SimulationBlock,
                                      TraditionalAlgorithmRK, ZeroGradientOutletBC)
    _announce("HIPTarget sod 256 x 100")
    block = SimulationBlock(1)
    block.set_block_boundaries([ZeroGradientOutletBC(0, 0), ZeroGradientOutletBC(0, 1)])
    block.set_discretisation_schemes({"w": LFWeno(5, formulation="Z", averaging=SimpleAverage([0, 1])),
                                      "rk": RungeKuttaLS(3)})
    params = {"gama": "1.4", "dt": "0.0002", "niter": "100", "block0np0": "256", "Delta0block0": "1.0/(block0np0-1)"}
    pr = P.sod(n=256, steps=100)
    with HIPTarget(TraditionalAlgorithmRK(block), parameters=params) as t:
        assert t.lib.path.endswith("libosb_hip.so")
        q = t.initialise(pr.q0).run().state()
    want, _ = run(oracle, pr)
    l2, li = rel_errors(q, want)
    assert l2.max() <= TOL and li.max() <= TOL


def test_tg_sym_known_answer(hip, oracle):
    """Reference pin: w at interior (8,8,8) after one step,
    tests/verification_apps/tg_sym/check_opensbli_output.py:31-34."""
    pr = P.tg_sym()
    a, b = _compare(hip, oracle, pr)
    w = a[3] / a[0]
    target = -2.2438363795512375e-08
    print(f"   w(8,8,8) = {w[8, 8, 8]!r}", flush=True)
    assert math.isclose(w[8, 8, 8], target, rel_tol=2e-8)


def test_tgv_teno6_reference_validation(hip):
    """Reference pin: apps/performance_benchmark/README.MD:46-47, TENO6 128^3, 500 iterations."""
    pr = P.tgv_teno6_validation()
    _announce(pr.name + " 500 steps (README validation)")
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(pr.steps)
        st = c.stats()
        d = c.tgv_diagnostics()
    print(f"   {st.iter}, {st.time:.5e}, " + ", ".join(f"{v:.5e}" for v in d), flush=True)
    ref = [2.65363e+05, 1.16991e+05, 1.79666e+06, 2.09715e+06]
    assert st.iter == 500
    assert f"{st.time:.5e}" == "8.46250e-01"
    for v, r in zip(d, ref):
        assert f"{v:.5e}" == f"{r:.5e}", (d, ref)


# --- HDF5 output and restart (libosb_io.so; opensbli/core/io_hdf5.py:165-250) ------
def test_hdf5_restart_on_the_gpu(hip, oracle, tmp_path):
    """The GPU state written in the OPS layout, read back by a restarted GPU
    context, continues bit for bit; its interior equals the oracle's file."""
    from opensbli_amd import io_hdf5 as io
    try:
        io.load_io()
    except FileNotFoundError as e:
        pytest.skip(str(e))
    for pr in (P.kelvin_helmholtz(n=64, steps=20), P.tgv(n=24, steps=6)):
        _announce(f"hdf5 restart {pr.name}")
        k = pr.steps // 2
        straight, st = run(hip, pr, pr.steps)
        paths = {}
        for name, lib in (("hip", hip), ("oracle", oracle)):
            paths[name] = str(tmp_path / f"{pr.name}_{name}.h5")
            with lib.context(pr.desc) as c:
                c.set_state(pr.q0)
                c.step(k)
                io.write_state(paths[name], c)
        qa, ia, ta = io.read_state(paths["hip"], pr.desc)
        qb, ib, tb = io.read_state(paths["oracle"], pr.desc)
        assert_exact_unless_fused(qa, qb, pr.desc)
        assert ia == ib == k and ta == tb
        with hip.context(pr.desc) as c:
            io.restart(c, paths["hip"])
            c.step(pr.steps - k)
            q = c.get_state()
            s2 = c.stats()
        assert np.array_equal(q, straight)
        assert s2.iter == st.iter and math.isclose(s2.time, st.time, rel_tol=1e-15)


def test_stage_by_stage_equals_step(hip, oracle):
    """osb_stage on the GPU: stage-by-stage equals osb_step (graphs) and the oracle's stages."""
    for pr in (P.kelvin_helmholtz(n=48, steps=2), P.tgv(n=16, steps=2)):
        _announce(f"stages {pr.name}")
        with hip.context(pr.desc) as a, hip.context(pr.desc) as b, oracle.context(pr.desc) as o:
            for c in (a, b, o):
                c.set_state(pr.q0)
            a.step(pr.steps)
            for _ in range(pr.steps):
                for s in range(3):
                    b.stage(s)
                    o.stage(s)
            qa, qb, qo = a.get_state(), b.get_state(), o.get_state()
        assert np.array_equal(qa, qb)
        assert_exact_unless_fused(qb, qo, pr.desc)


def test_probes_on_the_gpu(hip):
    pr = P.tgv(n=16, steps=2)
    _announce("probes tgv16")
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(2)
        q = c.get_state()
        pts = [(0, 0, 0), (15, 3, 7), (8, 15, 1)]
        v = c.probe(pts, [0, 4, 2])
    assert list(v) == [q[0, 0, 0, 0], q[4, 7, 3, 15], q[2, 1, 15, 8]]


# --- WENO filter after each full step (WENO_filter.py:406-536) ----------------------
def _filtered(pr, **kw):
    d = nat.ProblemDesc(**{**pr.desc.__dict__, "shock_filter": nat.FILTER_WENOZ, **kw})
    return P.Problem(pr.name + "_filtered", d, pr.q0, pr.steps)


@pytest.mark.parametrize("maker", [
    lambda: P.entropy_wave(n=64, profile="step", steps=20),                          # 2-D central, unfused
    lambda: P.entropy_wave(n=48, profile="step", steps=10, scheme="kgp"),            # 2-D KGP
    lambda: _filtered(P.tgv(n=16, steps=5, scheme="kgp")),                           # 3-D KGP, viscous
    lambda: _filtered(P.tgv(n=16, steps=5, minf=1.25, cfl=0.5, dt=0.0)),              # 3-D fused central, CFL
    lambda: _filtered(P.tg_sym(n=17, steps=3)),                                      # classic RK3, symmetry walls
    lambda: _filtered(P.sod(n=200, steps=20)),                                       # 1-D WENO base, simple average
    lambda: _filtered(P.kelvin_helmholtz(n=48, steps=10)),                           # 2-D WENO base, walls
])
def test_weno_filter_parity(hip, oracle, maker):
    pr = maker()
    a, b = _compare(hip, oracle, pr)
    if not fused_3d(pr.desc):
        assert np.array_equal(a, b)


def test_weno_filter_stages_and_graphs(hip, oracle, monkeypatch):
    """The filter runs after the last stage: stage-by-stage equals whole steps,
    and captured steps replayed equal eager launches."""
    pr = P.entropy_wave(n=32, profile="step", steps=3)
    _announce("filter stages / graph replay")
    out = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("OSB_GRAPH", graphs)
        with hip.context(pr.desc) as c:
            c.set_state(pr.q0)
            snaps = []
            for k in (1, 2, 1):
                c.step(k)
                snaps.append(c.get_state())
            out[graphs] = snaps
    with hip.context(pr.desc) as c, oracle.context(pr.desc) as o:
        for x in (c, o):
            x.set_state(pr.q0)
        for _ in range(4):
            for s in range(3):
                c.stage(s)
                o.stage(s)
        qs, qo = c.get_state(), o.get_state()
    for a, b in zip(out["1"], out["0"]):
        assert np.array_equal(a, b)
    assert np.array_equal(qs, out["0"][-1]) and np.array_equal(qs, qo)


# --- Carpenter boundary closure next to extrapolation / Dirichlet faces ------------
def _faces(pr, bc, values=None, name=""):
    d = nat.ProblemDesc(**{**pr.desc.__dict__, "bc": bc, "bc_value": values or {}})
    return P.Problem(pr.name + name, d, pr.q0, pr.steps)


_EXT3 = [[nat.BC_EXTRAPOLATION0, nat.BC_EXTRAPOLATION1], [nat.BC_PERIODIC] * 2, [nat.BC_EXTRAPOLATION1, nat.BC_EXTRAPOLATION0]]


def test_residual_closure_tgv(hip, oracle):
    _residual_compare(hip, oracle, _faces(P.tgv(n=20), _EXT3, name="_ext"))


@pytest.mark.parametrize("maker", [
    lambda: P.sod_filtered(n=200, steps=100),                                       # 1-D central + filter + Dirichlet
    lambda: _faces(P.tgv(n=16, steps=4), _EXT3, name="_ext"),                       # 3-D Feiereisen viscous, unfused
    lambda: _faces(P.tgv(n=16, steps=4, scheme="kgp"),
                   [[nat.BC_DIRICHLET] * 2, [nat.BC_PERIODIC] * 2, [nat.BC_PERIODIC] * 2],
                   {(0, 0): [1.0, 0.0, 0.0, 0.0, 179.0], (0, 1): [1.0, 0.0, 0.0, 0.0, 179.0]}, "_dir"),   # 3-D KGP
    lambda: _faces(P.tgv(n=16, steps=4, scheme="teno", minf=1.25, cfl=0.5, dt=0.0), _EXT3, name="_ext"),  # shock + viscous
    lambda: _faces(P.entropy_wave(n=40, profile="sine", steps=6, filtered=False),
                   [[nat.BC_EXTRAPOLATION1] * 2, [nat.BC_PERIODIC] * 2], name="_ext"),                  # 2-D inviscid
])
def test_closure_parity(hip, oracle, maker):
    pr = maker()
    a, b = _compare(hip, oracle, pr)
    if not fused_3d(pr.desc):
        assert np.array_equal(a, b)


# --- WENO3-Z filter with the Ducros hybrid selection (Central4-WENO5) -----------------
def _kgp2d_smooth(n=40, steps=12):
    gama = 1.4
    h = 2 * math.pi / n
    X, Y = np.meshgrid(h * np.arange(n), h * np.arange(n), indexing="xy")
    rho = 1.0 + 0.2 * np.sin(X) * np.cos(2 * Y)
    u, v = 0.6 * np.sin(X + 0.5) * np.cos(Y), -0.4 * np.cos(X) * np.sin(Y - 0.3)
    p = 1.0 + 0.1 * np.cos(X - Y)
    q = np.stack([rho, rho * u, rho * v, p / (gama - 1.0) + 0.5 * rho * (u * u + v * v)])[:, None]
    desc = nat.ProblemDesc(ndim=2, n=[n, n], delta=[h, h], conv_scheme=nat.CONV_CENTRAL_KGP, order=4,
                           bc=[[nat.BC_PERIODIC] * 2] * 2, gama=gama, dt=5e-3, shock_filter=nat.FILTER_WENOZ,
                           filter_order=3, filter_sensor=nat.SENSOR_DUCROS)
    return P.Problem("kgp2d_weno3_ducros", desc, q, steps)


@pytest.mark.parametrize("maker", [
    lambda: P.tgv_central4_weno5(n=24, steps=10),                                      # the benchmark's method, 3-D
    lambda: _kgp2d_smooth(),                                                            # 2-D hybrid
    lambda: P.Problem("entropy_o3", nat.ProblemDesc(**{**P.entropy_wave(n=48, steps=10).desc.__dict__, "filter_order": 3}),
                      P.entropy_wave(n=48, steps=10).q0, 10),                           # order 3, every point
])
def test_weno3_ducros_filter_parity(hip, oracle, maker):
    pr = maker()
    a, b = _compare(hip, oracle, pr)
    if not fused_3d(pr.desc):
        assert np.array_equal(a, b)


# --- error paths that stay on the host side of every kernel ---------------------------
@pytest.mark.parametrize("maker", [
    lambda: P.tgv(n=12, steps=1),
    lambda: P.tgv(n=8, scheme="teno", steps=1),
    lambda: P.entropy_wave(n=16, ndim=1, steps=1),
    lambda: P.entropy_wave(n=16, ndim=2, steps=1),
], ids=["tgv3d_h2", "tgv3d_h4", "wave1d_h4", "wave2d_h4"])
def test_periodic_halo_fill_one_pass(hip, maker, monkeypatch):
    """Every face periodic: the one-pass halo fill leaves each halo point,
    corners included, equal to the interior point at its index wrapped in
    every direction (numpy's wrap padding), and bit-equal to the
    direction-by-direction fill (OSB_BC_KERNEL=dir)."""
    pr = maker()
    nd = pr.desc.ndim
    assert all(pr.desc.bc[d][0] == nat.BC_PERIODIC and pr.desc.bc[d][1] == nat.BC_PERIODIC for d in range(nd))
    _announce(f"periodic halo fill {pr.name} n={list(pr.desc.n)}")
    out = {}
    for mode in ("one_pass", "dir"):
        monkeypatch.setenv("OSB_BC_KERNEL", mode)
        with hip.context(pr.desc) as c:
            shp = c.padded_shape()
            h = int(c.stats().halo)
            qp = np.random.default_rng(7).standard_normal(shp)
            c.set_state(qp, nat.LAYOUT_PADDED)
            c.apply_bcs()
            out[mode] = c.get_state(nat.LAYOUT_PADDED).reshape(shp)
    pads = [(0, 0)] + [(h, h) if 3 - a - 1 < nd else (0, 0) for a in range(3)]
    inner = tuple([slice(None)] + [slice(h, -h) if 3 - a - 1 < nd else slice(None) for a in range(3)])
    want = np.pad(qp[inner], pads, mode="wrap")
    assert np.array_equal(out["one_pass"], want)
    assert np.array_equal(out["dir"], want)


def _tgv_box(n, scheme="central", precision=nat.PREC_FP64, steps=2):
    """The TGV on a periodic box of n = [n0, n1, n2] points (2 pi per direction),
    anisotropic and not a multiple of the stage kernel's 16 x 4 x 4 tile."""
    pr = P.tgv(n=16, steps=steps, scheme=scheme, precision=precision)
    delta = [2 * math.pi / m for m in n]
    q = P._tgv_fields(tuple(n), tuple(delta), pr.desc.gama, pr.desc.minf)
    d = nat.ProblemDesc(**{**pr.desc.__dict__, "n": list(n), "delta": delta})
    return P.Problem(f"tgv_{scheme}_{n[0]}x{n[1]}x{n[2]}", d, q, steps)


@pytest.mark.parametrize("maker", [
    lambda: P.tgv(n=16, steps=2),
    lambda: P.tgv(n=16, steps=2, scheme="kgp"),
    lambda: P.tgv(n=16, steps=2, scheme="teno", minf=1.25),
    lambda: P.tgv(n=16, steps=2, precision=nat.PREC_MIXED),
    lambda: P.tgv(n=16, steps=2, scheme="kgp", precision=nat.PREC_MIXED),
    lambda: _tgv_box([20, 12, 10]),
    lambda: _tgv_box([20, 12, 10], scheme="kgp"),
    lambda: _tgv_box([20, 12, 10], precision=nat.PREC_MIXED),
    lambda: _tgv_box([20, 12, 10], scheme="kgp", precision=nat.PREC_MIXED),
], ids=["feiereisen", "kgp", "teno_viscous", "mixed_feiereisen", "mixed_kgp",
        "feiereisen_20x12x10", "kgp_20x12x10", "mixed_feiereisen_20x12x10", "mixed_kgp_20x12x10"])
def test_stage_kernel_halo_images(hip, maker, monkeypatch):
    """Every face periodic: the 3-D tile stage kernels (FP64 and mixed) write
    each updated point's wrapped halo images themselves, and the padded state after every
    stage (halos included) equals the one with the separate one-pass fill
    (OSB_BC_KERNEL=pass) and the per-direction fill (OSB_BC_KERNEL=dir)."""
    pr = maker()
    _announce(f"stage-kernel halo images {pr.name} n={list(pr.desc.n)}")
    out = {}
    for mode in ("in_stage", "pass", "dir"):
        monkeypatch.setenv("OSB_BC_KERNEL", mode)
        with hip.context(pr.desc) as c:
            c.set_state(pr.q0)
            snaps = []
            for _ in range(pr.steps):
                for s in range(3):
                    c.stage(s)
                    snaps.append(c.get_state(nat.LAYOUT_PADDED))
            c.step(2)
            snaps.append(c.get_state(nat.LAYOUT_PADDED))
            out[mode] = snaps
    for a, b, d in zip(out["in_stage"], out["pass"], out["dir"]):
        assert np.array_equal(a, b) and np.array_equal(a, d)


@pytest.mark.parametrize("scheme", ["central", "kgp"])
def test_tgv_anisotropic_ragged_tiles(hip, oracle, scheme):
    """A 20 x 12 x 10 box: partly filled stage-kernel tiles in every direction
    and distinct extents (an n0/n1/n2 mix-up would show), FP64 against the
    oracle and mixed precision against the FP64 oracle."""
    pr = _tgv_box([20, 12, 10], scheme=scheme, steps=6)
    _compare(hip, oracle, pr)
    mx = _tgv_box([20, 12, 10], scheme=scheme, precision=nat.PREC_MIXED, steps=6)
    _announce(f"mixed {mx.name}")
    a, _ = run(hip, mx)
    b, _ = run(oracle, pr)
    l2, _ = rel_errors(a, b)
    print(f"   mixed vs FP64 oracle rel L2 {np.array2string(l2, precision=2)}", flush=True)
    assert l2.max() <= 1e-5, l2


@pytest.mark.parametrize("tile", ["0", "1", "2"])
def test_fused_stage_tiles_against_oracle(hip, oracle, tile, monkeypatch):
    """Every tile shape the fused 3-D stage kernel offers (forced with
    OSB_FUSED_TILE) meets the oracle on a ragged anisotropic box (Feiereisen,
    KGP + Sutherland) and on the TENO5 + viscous path."""
    monkeypatch.setenv("OSB_FUSED_TILE", tile)
    ks = _tgv_box([20, 12, 10], scheme="kgp", steps=4)
    ks.desc.viscous = nat.VISC_SUTHERLAND
    ks.desc.minf = 1.25
    ks.q0 = P._tgv_fields((20, 12, 10), tuple(ks.desc.delta), ks.desc.gama, 1.25)
    for pr in (_tgv_box([20, 12, 10], steps=4), ks, P.tgv(n=16, steps=3, scheme="teno", minf=1.25)):
        with hip.context(pr.desc) as c:
            info = c.describe()
        assert info["tile_choice"] == "forced", info
        _compare(hip, oracle, pr)


@pytest.mark.parametrize("width", ["16", "32"])
def test_face_resid_width_bitwise(hip, oracle, width, monkeypatch):
    """Both widths of the fused face-residual tiles (forced with OSB_RESID_W)
    round exactly like the oracle: 2-D on a box that fills neither width's
    tiles (WENO-Z, KH; TENO5 on a Riemann problem) and 3-D inviscid WENO-Z."""
    monkeypatch.setenv("OSB_RESID_W", width)
    kh = P.kelvin_helmholtz(n=40, steps=6)
    rm = P.riemann_2d(n=36, steps=6)
    rm.desc.conv_scheme = nat.CONV_TENO
    rm.name += "_teno5"
    w3 = P.tgv(n=20, steps=4, scheme="weno", minf=1.25, cfl=0.5, dt=0.0)
    w3.desc.viscous = nat.VISC_NONE
    w3.name += "_inviscid"
    for pr in (kh, rm, w3):
        with hip.context(pr.desc) as c:
            c.set_state(pr.q0)
            c.step(1)
            info = c.describe()
        assert info["resid_w"] == int(width) and info["resid_w_choice"] == "forced", info
        a, _ = run(hip, pr)
        b, _ = run(oracle, pr)
        assert np.array_equal(a, b), pr.name


def test_face_resid_width_autotuned(hip):
    """At the first step the face-residual widths are timed on the context's
    state and the fastest kept; the choice does not change the results."""
    pr = P.kelvin_helmholtz(n=96, steps=3)
    _announce("face-resid width autotune kh96")
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(3)
        q = c.get_state()
        info = c.describe()
    print("  ", info, flush=True)
    us = info["resid_w_us"]
    assert info["resid_w_choice"] == "autotuned" and len(us) == 2 and min(us) > 0
    assert kept_fastest(info["resid_w"], [32, 16], us)
    for w in ("16", "32"):
        os.environ["OSB_RESID_W"] = w
        try:
            forced, _ = run(hip, pr)
        finally:
            del os.environ["OSB_RESID_W"]
        assert np.array_equal(forced, q), w


def test_fused_stage_tile_autotuned(hip):
    """At the first step the tile shapes are timed on the context's own state and
    the fastest kept; the choice does not change the results."""
    pr = P.tgv(n=48, steps=2)
    _announce("tile autotune tgv48")
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        assert c.describe()["tile_choice"] == "default"
        c.step(2)
        q = c.get_state()
        info = c.describe()
    print("  ", info, flush=True)
    us = info["tile_us"]
    assert info["tile_choice"] == "autotuned" and len(us) == 3 and min(us) > 0
    tiles = [[16, 4, 4], [16, 8, 4], [8, 8, 8]]
    assert kept_fastest(info["tile"], tiles, us)
    for t in range(3):
        os.environ["OSB_FUSED_TILE"] = str(t)
        try:
            forced, _ = run(hip, pr)
        finally:
            del os.environ["OSB_FUSED_TILE"]
        assert np.array_equal(forced, q), t


def test_out_of_memory_is_an_error_not_a_crash(hip):
    """A grid far beyond HBM fails context creation with OSB_E_NOMEM (every
    partial allocation released); the library keeps working afterwards, and
    zero / negative step counts behave as documented."""
    # descriptor only (no host state): 1600^3 points, ~3 x 166 GB of state buffers
    big = nat.ProblemDesc(**{**P.tgv(n=16).desc.__dict__, "n": [1600] * 3, "delta": [2 * math.pi / 1600] * 3})
    _announce("oversized context")
    assert hip.estimate_bytes(big) > 400 * 2 ** 30
    with pytest.raises(nat.OSBError) as e:
        hip.context(big)
    assert e.value.code == nat.E_NOMEM
    # beyond 2^32 points per variable (the kernels' 32-bit point offsets): refused before any allocation
    huge = nat.ProblemDesc(**{**big.__dict__, "n": [2048] * 3, "delta": [2 * math.pi / 2048] * 3})
    with pytest.raises(nat.OSBError) as e:
        hip.context(huge)
    assert e.value.code == nat.E_UNSUPPORTED
    pr = P.kelvin_helmholtz(n=32, steps=2)
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(0)
        assert c.stats().iter == 0
        with pytest.raises(nat.OSBError):
            c.step(-1)
        c.step(2)
        q = c.get_state()
    want, _ = run(hip, pr)
    assert np.array_equal(q, want)


# --- long horizons: a one-ulp difference would have grown by now (bitwise paths), and
# the fused 3-D path stays within the FP64 bar over hundreds of steps --------------
@pytest.mark.parametrize("maker", [lambda: P.kelvin_helmholtz(n=128, steps=1000),   # WENO-Z, fused faces, walls
                                   lambda: P.tgv(n=32, steps=300),                    # fused central + viscous
                                   lambda: P.tgv_central4_weno5(n=24, steps=200)])    # fused KGP + hybrid filter
def test_long_horizon(hip, oracle, maker):
    pr = maker()
    a, b = _compare(hip, oracle, pr)
    if not fused_3d(pr.desc):
        assert np.array_equal(a, b)


# --- HLLC / HLLC-LM -----------------------------------------------------------------
def _hllc(pr, flux):
    d = nat.ProblemDesc(**{**pr.desc.__dict__, "flux_type": flux})
    return P.Problem(pr.name + f"_hllc{flux}", d, pr.q0, pr.steps)


@pytest.mark.parametrize("flux", [nat.FLUX_HLLC, nat.FLUX_HLLC_LM])
@pytest.mark.parametrize("maker", [lambda: P.sod(n=400, steps=100), lambda: P.riemann_2d(n=48, steps=20),
                                   lambda: P.kelvin_helmholtz(n=48, steps=20),
                                   lambda: P.tgv(n=16, steps=5, scheme="weno", minf=1.25, cfl=0.5, dt=0.0)])
def test_hllc_parity(hip, oracle, maker, flux):
    pr = _hllc(maker(), flux)
    a, b = _compare(hip, oracle, pr)
    if not fused_3d(pr.desc):
        assert np.array_equal(a, b)
