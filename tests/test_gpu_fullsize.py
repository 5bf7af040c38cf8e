"""The headline configurations at their stated sizes, HIP library against the
CPU oracle: the viscous Taylor-Green vortex at 128^3 for its 200 steps (FP64
at the FP64 bar, mixed precision with its error against the FP64 oracle
stated), and the supersonic TENO5 + Sutherland TGV at 256^3 with the device
CFL time step (first steps of the run; FP64 and mixed)."""
import time

import numpy as np
import pytest

from opensbli_amd import native as nat, problems as P
from helpers import fp64_bar, rel_errors, run

pytestmark = pytest.mark.gpu

MIXED_BAR = 1e-5


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)


def _mixed(pr):
    d = nat.ProblemDesc(**{**pr.desc.__dict__, "precision": nat.PREC_MIXED})
    return P.Problem(pr.name + "_mixed", d, pr.q0, pr.steps)


@pytest.fixture(scope="module")
def tgv128(oracle):
    pr = P.tgv(n=128, steps=200)
    _announce(f"oracle {pr.name} x {pr.steps} steps")
    t0 = time.time()
    b, sb = run(oracle, pr)
    print(f"   oracle {time.time() - t0:.1f} s", flush=True)
    return pr, b, sb


def test_tgv_central_128_full_config_fp64(hip, oracle, tgv128):
    """Config 2: Feiereisen + constant viscosity, RK3-LS, 128^3, 200 steps."""
    pr, b, sb = tgv128
    _announce(f"hip {pr.name} x {pr.steps} steps FP64")
    a, sa = run(hip, pr)
    l2, li = rel_errors(a, b)
    print(f"   rel L2 {np.array2string(l2, precision=2)}  rel Linf max {li.max():.3e}", flush=True)
    assert np.all(np.isfinite(a))
    fp64_bar(a, b, pr.desc, oracle, pr)
    assert sa.iter == sb.iter == 200 and abs(sa.time - sb.time) <= 1e-12 * sb.time


def test_tgv_central_128_full_config_mixed(hip, tgv128):
    """Config 2, mixed precision (FP32 derivatives over the FP64 state): error
    against the FP64 oracle after the 200 steps, per conservative variable."""
    pr, b, _ = tgv128
    mx = _mixed(pr)
    _announce(f"hip {mx.name} x {mx.steps} steps")
    a, _ = run(hip, mx)
    l2, li = rel_errors(a, b)
    whole = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"   mixed vs FP64 oracle: state rel L2 {whole:.3e}; per variable rel L2 {np.array2string(l2, precision=2)}, "
          f"rel Linf max {li.max():.3e}", flush=True)
    assert l2.max() <= MIXED_BAR, l2


@pytest.fixture(scope="module")
def teno256(oracle):
    pr = P.tgv(n=256, steps=3, scheme="teno", minf=1.25, cfl=0.5, dt=0.0)
    _announce(f"oracle {pr.name} x {pr.steps} CFL steps")
    t0 = time.time()
    b, sb = run(oracle, pr)
    print(f"   oracle {time.time() - t0:.1f} s, dt {sb.last_dt:.6e}", flush=True)
    return pr, b, sb


def test_tgv_teno5_256_cfl_fp64(hip, oracle, teno256):
    """Config 3: TENO5 (C_T = 1e-5) with Roe averaging, LLF, Sutherland viscous
    terms, RK3-LS and dt from the CFL = 0.5 device reduction, 256^3."""
    pr, b, sb = teno256
    _announce(f"hip {pr.name} x {pr.steps} CFL steps FP64")
    a, sa = run(hip, pr)
    l2, li = rel_errors(a, b)
    print(f"   rel L2 {np.array2string(l2, precision=2)}  rel Linf max {li.max():.3e}; dt {sa.last_dt:.6e}", flush=True)
    assert np.all(np.isfinite(a))
    bound = fp64_bar(a, b, pr.desc, oracle, pr)
    assert sa.iter == sb.iter and abs(sa.time - sb.time) <= max(1e-12, bound) * sb.time


def test_tgv_teno5_256_cfl_mixed(hip, teno256):
    """Config 3, mixed precision (FP32 characteristic and derivative work; FP64
    state, RK, smoothness indicators and TENO cut-off): error against the FP64
    oracle, per conservative variable."""
    pr, b, _ = teno256
    mx = _mixed(pr)
    _announce(f"hip {mx.name} x {mx.steps} CFL steps")
    a, _ = run(hip, mx)
    l2, li = rel_errors(a, b)
    whole = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"   mixed vs FP64 oracle: state rel L2 {whole:.3e}; per variable rel L2 {np.array2string(l2, precision=2)}, "
          f"rel Linf max {li.max():.3e}", flush=True)
    assert l2.max() <= MIXED_BAR, l2



@pytest.mark.parametrize("n,order", [(96, "XCD ranges"), (80, "XCD ranges"), (72, "linear"), (192, "XCD ranges")])
def test_tgv_central_tile_orders_at_size(hip, oracle, n, order, monkeypatch):
    """The stage kernels' tile orders other than the 128^3 supertile walk, on
    TGV cubes whose 16 x 4 x 4 tile counts (forced by OSB_FUSED_TILE) do not
    take supertiles: XCD ranges of linear tiles (96^3, 80^3, 192^3; 80^3 with
    ragged x tiles) and the plain linear order (72^3: 1620 tiles, not a
    multiple of 8).  After 4 steps, FP64 against the oracle at the FP64 bar and
    mixed precision within its bar of the FP64 oracle."""
    monkeypatch.setenv("OSB_FUSED_TILE", "0")
    pr = P.tgv(n=n, steps=4)
    _announce(f"tile order '{order}': oracle {pr.name} x {pr.steps} steps")
    b, sb = run(oracle, pr)
    _announce(f"tile order '{order}': hip {pr.name} FP64")
    a, sa = run(hip, pr)
    assert np.all(np.isfinite(a))
    fp64_bar(a, b, pr.desc, oracle, pr)
    assert sa.iter == sb.iter == 4
    mx = _mixed(pr)
    _announce(f"tile order '{order}': hip {mx.name}")
    m, _ = run(hip, mx)
    l2, _ = rel_errors(m, b)
    assert l2.max() <= MIXED_BAR, l2
