"""The random configuration matrix runs on the oracle: every case is a
descriptor both libraries accept, and three steps stay finite."""
import numpy as np
import pytest

from config_matrix import CASES, make_case
from helpers import run


def test_matrix_cases_run_on_the_oracle(oracle):
    kinds = set()
    for seed in CASES:
        pr = make_case(seed)
        q, st = run(oracle, pr)
        assert st.iter == 3 and np.all(np.isfinite(q)), pr.name
        d = pr.desc
        kinds.add((d.ndim, d.conv_scheme, d.shock_filter, d.filter_sensor))
    assert len(kinds) >= 15, kinds
