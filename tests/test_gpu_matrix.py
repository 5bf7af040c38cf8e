"""Parity sweep over seeded random configurations (tests/config_matrix.py):
each runs three steps on the GPU and the oracle, bitwise equal."""
import numpy as np
import pytest

from config_matrix import CASES, make_case
from helpers import run

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
        if not (np.array_equal(a, b) and sa.time == sb.time):
            bad.append((pr.name, float(np.abs(a - b).max())))
    assert not bad, bad
