"""Batched ensembles and the serving pool on the GPU.

Bar: every instance advanced inside a batch, or inside the pool next to other
solves, is bitwise identical to the same instance stepped alone (the batched
launches run the same per-point code), and the FP64 ones match the oracle.
Mixed precision is checked against the FP64 oracle at the stated 1e-5
relative L2."""
import os

import numpy as np
import pytest

from opensbli_amd import native as nat, problems as P
from helpers import kept_fastest, rel_errors, run

pytestmark = pytest.mark.gpu


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)


def _tgv_box(n, steps=5, dt_n=16):
    """TGV fields on an n[0] x n[1] x n[2] box, with the time step of the
    dt_n^3 cube"""
    import math
    pr = P.tgv(n=dt_n, steps=steps)
    delta = [2 * math.pi / m for m in n]
    q = P._tgv_fields(tuple(n), tuple(delta), pr.desc.gama, pr.desc.minf)
    d = nat.ProblemDesc(**{**pr.desc.__dict__, "n": list(n), "delta": delta})
    return P.Problem(f"tgv_central_{n[0]}x{n[1]}x{n[2]}", d, q, steps)


def _ensemble():
    """Instances of several groups and sizes, and two that run unbatched."""
    probs = []
    for n, prec in ((32, nat.PREC_FP64), (48, nat.PREC_FP64), (40, nat.PREC_MIXED), (64, nat.PREC_MIXED)):
        pr = P.kelvin_helmholtz(n=n, steps=5, precision=prec)
        pr.name += f"_p{prec}"
        probs.append(pr)
    probs.append(P.sod(n=128, steps=5))
    probs.append(P.sod(n=96, steps=5, boundaries="dirichlet"))
    probs.append(P.sod(n=80, steps=5, boundaries="linear"))
    probs.append(P.riemann_2d(n=36, steps=5, boundaries="mixed"))
    kg = P.kelvin_helmholtz(n=40, steps=5)
    kg.desc.flux_type = nat.FLUX_GLF                  # GLF next to LLF items of the same group
    kg.name += "_glf"
    probs.append(kg)
    kgm = P.kelvin_helmholtz(n=44, steps=5, precision=nat.PREC_MIXED)
    kgm.desc.flux_type = nat.FLUX_GLF
    kgm.name += "_glf_mixed"
    probs.append(kgm)
    r = P.riemann_2d(n=40, steps=5)
    r.desc.cfl, r.desc.dt = 0.4, 0.0                  # device CFL inside the batch
    probs.append(r)
    k3 = P.kelvin_helmholtz(n=36, steps=5)
    k3.desc.time_scheme = nat.TIME_RK3                # classic RK3 group
    k3.name += "_rk3"
    probs.append(k3)
    from opensbli_amd.surface import RungeKuttaLS
    k4 = P.kelvin_helmholtz(n=36, steps=5)
    ssp = RungeKuttaLS(3, stages=4, formulation="SSP")   # even stage count
    k4.desc.rk_a, k4.desc.rk_b = ssp.rk_a, ssp.rk_b
    k4.name += "_ssp4"
    probs.append(k4)
    w3 = P.tgv(n=16, steps=5, scheme="weno", minf=1.25, cfl=0.5, dt=0.0)
    w3.desc.viscous = nat.VISC_NONE                   # 3-D shock-only group
    w3.name += "_inviscid"
    probs.append(w3)
    # 3-D central instances (FP64, RK-LS, every face periodic): batched central
    # groups by split form and viscosity, any sizes, CFL next to constant dt
    probs.append(P.tgv(n=16, steps=5))
    probs.append(P.tgv(n=24, steps=5))
    probs.append(_tgv_box([20, 12, 10], steps=5))
    probs.append(P.tgv(n=16, steps=5, scheme="kgp"))
    ks = P.tgv(n=20, steps=5, scheme="kgp", minf=1.25)
    ks.desc.viscous = nat.VISC_SUTHERLAND
    ks.name += "_sutherland"
    probs.append(ks)
    sk = P.tgv(n=16, steps=5)
    sk.desc.conv_scheme = nat.CONV_CENTRAL_SKEW
    sk.name += "_skew"
    probs.append(sk)
    cf = P.tgv(n=16, steps=5, minf=1.25, cfl=0.5, dt=0.0)
    cf.name += "_cfl"
    probs.append(cf)
    probs.append(P.tgv(n=16, steps=5, precision=nat.PREC_MIXED))     # mixed central: stepped alone
    cl = P.tgv(n=16, steps=5)
    cl.desc.bc = [[nat.BC_EXTRAPOLATION0, nat.BC_EXTRAPOLATION1], [nat.BC_PERIODIC] * 2, [nat.BC_PERIODIC] * 2]
    cl.name += "_closure"                                              # closure faces: stepped alone
    probs.append(cl)
    probs.append(P.tgv(n=16, steps=5, scheme="teno", minf=1.25, cfl=0.5, dt=0.0))
    sf = P.sod(n=112, steps=5)
    sf.desc.shock_filter = nat.FILTER_WENOZ           # filtered contexts are stepped alone
    sf.name += "_filtered"
    probs.append(sf)
    return probs


@pytest.mark.parametrize("graphs", ["1", "0"])
def test_batch_bitwise_equals_single(hip, graphs, monkeypatch):
    monkeypatch.setenv("OSB_BATCH_GRAPH", graphs)
    probs = _ensemble()
    _announce(f"batch of {len(probs)} graphs={graphs}")
    with hip.batch([p.desc for p in probs]) as b:
        for i, p in enumerate(probs):
            b.set_state(i, p.q0)
        b.step(3)
        b.step(2)
        b.sync()
        info = b.describe()
        print("  ", info, flush=True)
        got = [b.get_state(i) for i in range(len(probs))]
        stats = [b.stats(i) for i in range(len(probs))]
    assert info["solo"] == 4 and len(info["groups"]) >= 8
    central = [g for g in info["groups"] if g["kind"] == "central"]
    assert len(central) == 4 and sum(g["members"] for g in central) == 7
    fused = [g for g in info["groups"] if g["kind"] == "shock" and g["ndim"] >= 2 and "resid_w" in g["tuning"]]
    assert fused
    for g in fused:
        # the fused shock groups' face-residual width was timed likewise
        tu = g["tuning"]
        assert kept_fastest(tu["resid_w"], [32, 16], tu["resid_w_us"]) and min(tu["resid_w_us"]) > 0, tu
    for g in central:
        # each central group's tile shape was timed on its own members before the first step
        tu = g["tuning"]
        assert tu["tile"] in ([16, 4, 4], [16, 8, 4], [8, 8, 8]) and min(tu["tile_us"]) > 0, tu
        assert kept_fastest(tu["tile"], [[16, 4, 4], [16, 8, 4], [8, 8, 8]], tu["tile_us"]), tu
    for p, q, st in zip(probs, got, stats):
        _announce(f"  compare {p.name}")
        want, wst = run(hip, p, steps=5)
        assert np.array_equal(q, want), p.name
        assert st.iter == wst.iter == 5
        assert st.time == wst.time, (p.name, st.time, wst.time)


@pytest.mark.parametrize("tile", [0, 1, 2])
def test_batch_central_tile_orders_bitwise(hip, tile, monkeypatch):
    """Batched 3-D central stages at each tile shape (forced by OSB_FUSED_TILE),
    with instances whose tile counts take the supertile walk (128 x 32 x 16 at
    every shape, 128 x 16 x 8 at 16 x 4 x 4) next to ones that walk tiles
    linearly: each instance bitwise equal to itself stepped alone."""
    monkeypatch.setenv("OSB_FUSED_TILE", str(tile))
    probs = [_tgv_box([128, 32, 16], steps=3, dt_n=128), _tgv_box([128, 16, 8], steps=3, dt_n=128),
             P.tgv(n=16, steps=3), _tgv_box([48, 20, 12], steps=3, dt_n=48)]
    _announce(f"central batch, tile {tile}")
    with hip.batch([p.desc for p in probs]) as b:
        for i, p in enumerate(probs):
            b.set_state(i, p.q0)
        b.step(3)
        b.sync()
        info = b.describe()
        got = [b.get_state(i) for i in range(len(probs))]
    assert all(np.all(np.isfinite(q)) for q in got)
    central = [g for g in info["groups"] if g["kind"] == "central"]
    assert len(central) == 1 and central[0]["members"] == len(probs), info
    assert central[0]["tuning"]["tile"] == [[16, 4, 4], [16, 8, 4], [8, 8, 8]][tile], info
    for p, q in zip(probs, got):
        _announce(f"  compare {p.name}")
        want, _ = run(hip, p, steps=3)
        assert np.array_equal(q, want), p.name


@pytest.mark.parametrize("width", [16, 32])
def test_batch_face_resid_width_bitwise(hip, width, monkeypatch):
    """Batched shock groups at each face-residual width (forced by
    OSB_RESID_W), FP64 and mixed, on boxes that fill neither width's tiles:
    each instance bitwise equal to itself stepped alone at that width."""
    monkeypatch.setenv("OSB_RESID_W", str(width))
    probs = [P.kelvin_helmholtz(n=40, steps=4), P.kelvin_helmholtz(n=72, steps=4), P.riemann_2d(n=36, steps=4),
             P.kelvin_helmholtz(n=44, steps=4, precision=nat.PREC_MIXED),
             P.kelvin_helmholtz(n=68, steps=4, precision=nat.PREC_MIXED)]
    _announce(f"shock batch, width {width}")
    with hip.batch([p.desc for p in probs]) as b:
        for i, p in enumerate(probs):
            b.set_state(i, p.q0)
        b.step(4)
        b.sync()
        info = b.describe()
        got = [b.get_state(i) for i in range(len(probs))]
    fused = [g for g in info["groups"] if g["kind"] == "shock" and "resid_w" in g["tuning"]]
    assert len(fused) >= 2 and all(g["tuning"]["resid_w"] == width for g in fused), info
    for p, q in zip(probs, got):
        _announce(f"  compare {p.name}")
        want, _ = run(hip, p, steps=4)
        assert np.array_equal(q, want), p.name


def test_batch_fp64_matches_oracle(hip, oracle):
    probs = [P.kelvin_helmholtz(n=48, steps=20), P.riemann_2d(n=48, steps=20), P.sod(n=200, steps=20),
             P.riemann_2d(n=40, steps=20, boundaries="mixed"), P.sod(n=160, steps=20, boundaries="dirichlet")]
    _announce("batch vs oracle")
    with hip.batch([p.desc for p in probs]) as b:
        for i, p in enumerate(probs):
            b.set_state(i, p.q0)
        b.step(20)
        b.sync()
        got = [b.get_state(i) for i in range(len(probs))]
    for p, q in zip(probs, got):
        want, _ = run(oracle, p, steps=20)
        l2, li = rel_errors(q, want)
        assert l2.max() <= 1e-11 and li.max() <= 1e-11, (p.name, l2, li)


def test_mixed_precision_error_vs_fp64_oracle(hip, oracle):
    """FP32 flux / derivative evaluation over the FP64 state: relative L2
    against the FP64 oracle, per conservative variable: KH 128^2 x 100 steps,
    TENO5 + viscous TGV 24^3 x 20 (CFL), Sod 1024 x 1000, the central
    viscous TGV 32^3 x 50 (Feiereisen and KGP) and KH 128^2 with GLF splitting
    (the stated bar is 1e-5)."""
    cases = [(P.kelvin_helmholtz(n=128, steps=100, precision=nat.PREC_MIXED), 100),
             (P.tgv(n=24, steps=20, scheme="teno", minf=1.25, cfl=0.5, dt=0.0, precision=nat.PREC_MIXED), 20),
             (P.sod(n=1024, steps=1000, precision=nat.PREC_MIXED), 1000),
             (P.tgv(n=32, steps=50, precision=nat.PREC_MIXED), 50),
             (P.tgv(n=32, steps=50, scheme="kgp", precision=nat.PREC_MIXED), 50)]
    kg = P.kelvin_helmholtz(n=128, steps=100, precision=nat.PREC_MIXED)
    kg.desc.flux_type = nat.FLUX_GLF
    kg.name += "_glf"
    cases.append((kg, 100))
    for pr, steps in cases:
        _announce(f"mixed {pr.name} {steps} steps")
        q, _ = run(hip, pr, steps=steps)
        ref_pr = P.Problem(pr.name, nat.ProblemDesc(**{**pr.desc.__dict__, "precision": nat.PREC_FP64}), pr.q0, steps)
        want, _ = run(oracle, ref_pr, steps=steps)
        l2, li = rel_errors(q, want)
        whole = np.linalg.norm(q - want) / np.linalg.norm(want)
        print(f"   mixed vs FP64 oracle: state rel L2 {whole:.3e}; per variable rel L2 "
              f"{np.array2string(l2, precision=2)}, rel Linf max {li.max():.3e}", flush=True)
        assert l2.max() <= 1e-5, l2


@pytest.mark.parametrize("policy", ["fifo", "sjf"])
def test_pool_results_same_alone_and_in_the_mix(hip, policy):
    """Requests of both kinds through the pool with a budget that forces
    queueing and joining mid-run; each result equals the same solve alone
    (arrival-order and least-work-first admission)."""
    from opensbli_amd import serving as S
    reqs = S.make_stream(count=10, rate=400.0, seed=7, riemann_sizes=(32, 48, 64), tgv_sizes=(16, 24),
                         steps_range=(10, 40))
    budget = max(hip.estimate_bytes(r.problem.desc) for r in reqs) * 3
    _announce(f"pool: {len(reqs)} requests, budget {budget} B")
    rep = S.serve(hip, reqs, budget, max_solves=4, nstreams=2, chunk=7, keep_results=True, policy=policy)
    print(f"   p50 {rep.p50 * 1e3:.1f} ms p99 {rep.p99 * 1e3:.1f} ms, peak running {rep.peak_running}", flush=True)
    assert len(rep.requests) == len(reqs)
    assert rep.peak_running >= 2
    for r in rep.requests:
        want, st = run(hip, r.problem, steps=r.steps)
        assert r.iters == r.steps == st.iter
        assert np.array_equal(r.result, want), r.problem.name


@pytest.mark.parametrize("maker", [lambda: P.sod(n=256, steps=7), lambda: P.kelvin_helmholtz(n=48, steps=7),
                                   lambda: P.tgv(n=16, steps=7),
                                   lambda: P.tgv(n=16, steps=7, scheme="teno", minf=1.25, cfl=0.5, dt=0.0)])
def test_graph_replay_equals_eager(hip, maker, monkeypatch):
    """Captured time steps replayed several times, with eager work and
    synchronisation between the replays, against the eager launches."""
    pr = maker()
    _announce(f"graph replay {pr.name}")
    out = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("OSB_GRAPH", graphs)
        with hip.context(pr.desc) as c:
            c.set_state(pr.q0)
            snaps = []
            for k in (1, 2, 1, 3):
                c.step(k)
                snaps.append(c.get_state())      # eager gather + sync between replays
            out[graphs] = (snaps, c.stats().time)
    for a, b in zip(out["1"][0], out["0"][0]):
        assert np.array_equal(a, b)
    assert out["1"][1] == out["0"][1]


def test_device_state_roundtrip_with_torch(hip):
    """osb_set_state_device / osb_get_state_device with torch tensors as glue."""
    import torch
    pr = P.kelvin_helmholtz(n=48, steps=4)
    _announce("device state transfer")
    src = torch.from_numpy(pr.q0.copy()).cuda()
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    with hip.context(pr.desc) as c:
        c.set_state_device(src.data_ptr())
        c.step(pr.steps)
        c.get_state_device(dst.data_ptr())
        host = c.get_state()
    want, _ = run(hip, pr)
    assert np.array_equal(dst.cpu().numpy(), want) and np.array_equal(host, want)
