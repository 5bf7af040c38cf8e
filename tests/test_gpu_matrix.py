"""Parity sweep over seeded random configurations (tests/config_matrix.py):
each runs three steps on the GPU and the oracle, bitwise equal (the fused 3-D
stage kernel's cases: within the FP64 bar, relative L2 and L-inf <= 1e-11, or
10x the oracle's own roundoff spread where a threshold selection makes that
larger -- helpers.fp64_bar)."""
import numpy as np
import pytest

from config_matrix import CASES, make_case
from helpers import fp64_bar, fused_3d, run

pytestmark = pytest.mark.gpu


def test_matrix_bitwise(hip, oracle):
    bad = []
    for seed in CASES:
        pr = make_case(seed)
        d = pr.desc
        print(f"[gpu-case] {pr.name} ndim={d.ndim} n={list(d.n)[:d.ndim]} conv={d.conv_scheme}/{d.order} "
              f"visc={d.viscous} bc={d.bc} flux={d.flux_type} filter={d.shock_filter}/{d.filter_order}/{d.filter_sensor} "
              f"time={d.time_scheme}/{len(d.rk_a)} cfl={d.cfl}", flush=True)
        a, sa = run(hip, pr)
        b, sb = run(oracle, pr)
        if fused_3d(d):
            try:
                bound = fp64_bar(a, b, d, oracle, pr)
                # with CFL stepping dt follows the state, so the time agrees to the same bar
                ok = abs(sa.time - sb.time) <= max(1e-12, bound) * abs(sb.time)
            except AssertionError:
                ok = False
        else:
            ok = np.array_equal(a, b) and sa.time == sb.time
        if not ok:
            bad.append((pr.name, float(np.abs(a - b).max())))
    assert not bad, bad
