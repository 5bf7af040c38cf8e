"""Shared helpers for the parity tests."""
import numpy as np


def rel_errors(a, b):
    """Per-variable relative L2 and L-inf of a against reference b (arrays shaped (nv, ...))."""
    l2, linf = [], []
    for v in range(b.shape[0]):
        ref = b[v].ravel()
        d = a[v].ravel() - ref
        nref2 = np.linalg.norm(ref)
        nrefi = np.max(np.abs(ref))
        l2.append(np.linalg.norm(d) / (nref2 if nref2 > 0 else 1.0))
        linf.append(np.max(np.abs(d)) / (nrefi if nrefi > 0 else 1.0))
    return np.array(l2), np.array(linf)


def run(lib, problem, steps=None, state=None):
    with lib.context(problem.desc) as c:
        c.set_state(problem.q0 if state is None else state)
        c.step(problem.steps if steps is None else steps)
        return c.get_state(), c.stats()
