"""Failure handling and initial-condition ingestion on the GPU: the NaN check
(ops_NaNcheck, helperfunctions.py:201-214) through libosb_hip.so, a serving
stream in which one request goes non-finite while the others still finish
bit-equal to their solves alone, and HIPTarget.initialise() driving the
scripts' GridBasedInitialisation equations (gridbasedinit.py:16-86) into
libosb_hip.so against the oracle.  Non-finite values come from the inputs
(a NaN in the uploaded state, a negative pressure); every kernel runs to
its normal end."""
import math

import numpy as np
import pytest

from opensbli_amd import native as nat, problems as P
from helpers import rel_errors, run

pytestmark = pytest.mark.gpu


def _announce(name):
    print(f"[gpu-case] {name}", flush=True)


def test_nan_check_on_the_gpu(hip):
    """osb_nan_check: OK on a finite state, OSB_E_NONFINITE after a NaN in
    the uploaded rho, and after a step from a state with negative pressure
    (the sound speed is NaN there)."""
    pr = P.kelvin_helmholtz(n=32, steps=2)
    _announce("nan check kh32")
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(2)
        assert c.nan_check() == nat.OK
        q = pr.q0.copy()
        q[0, 0, 5, 7] = np.nan
        c.set_state(q)
        assert c.nan_check() == nat.E_NONFINITE
    bad = pr.q0.copy()
    bad[3] *= -1.0                               # rhoE < 0: negative pressure everywhere
    with hip.context(pr.desc) as c:
        c.set_state(bad)
        c.step(1)
        assert c.nan_check() == nat.E_NONFINITE
    # the library keeps working afterwards
    with hip.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(2)
        assert c.nan_check() == nat.OK
        assert np.array_equal(c.get_state(), run(hip, pr)[0])


def test_serving_one_poisoned_request_fails_alone(hip):
    """A request whose initial state has a negative pressure quadrant goes
    non-finite; the serving loop reports it failed, and every other request
    (sharing its batched launches) finishes bit-equal to its solve alone."""
    from opensbli_amd import serving as S
    reqs = S.make_stream(count=8, rate=400.0, seed=11, riemann_sizes=(32, 48), tgv_sizes=(16,), steps_range=(10, 30))
    riemann = [r for r in reqs if "riemann" in r.problem.name]
    assert riemann, "the stream holds Riemann requests"
    victim = riemann[0]
    q = victim.problem.q0.copy()
    n = q.shape[-1]
    # rhoE = 0.2 * kinetic energy in the upper-right quadrant: p < 0 there
    ke = 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]
    sl = (slice(None), slice(n // 2, None), slice(n // 2, None))
    q[3][sl] = 0.2 * ke[sl] - 0.05
    victim.problem = P.Problem(victim.problem.name + "_poisoned", victim.problem.desc, q, victim.steps)
    budget = max(hip.estimate_bytes(r.problem.desc) for r in reqs) * 3
    _announce(f"serving with a poisoned request ({victim.problem.name}), {len(reqs)} requests")
    rep = S.serve(hip, reqs, budget, max_solves=4, nstreams=2, chunk=5, keep_results=True)
    assert len(rep.requests) == len(reqs)
    assert [r.rid for r in rep.failed] == [victim.rid]
    for r in rep.requests:
        if r.rid == victim.rid:
            assert not np.all(np.isfinite(r.result))
            continue
        want, st = run(hip, r.problem, steps=r.steps)
        assert r.finite and r.iters == r.steps == st.iter
        assert np.array_equal(r.result, want), r.problem.name


_SOD_PARAMS = {"gama": "1.4", "dt": "0.0002", "niter": "60", "block0np0": "512", "Delta0block0": "1.0/(block0np0-1)"}


def test_hip_target_initialise_from_script_equations(hip, oracle):
    """HIPTarget.initialise() evaluates the Sod and TGV scripts' initial-condition
    equations (GridBasedInitialisation strings) on the host and uploads them to
    libosb_hip.so; the runs equal the oracle's from the prepared states."""
    from test_target import _SOD_INIT, _TGV_INIT, _script_block
    from opensbli_amd.code_generation.hip import HIPTarget
    from opensbli_amd.code_generation.hip.target import initial_fields
    from opensbli_amd.surface import (LFWeno, NS_Split, PeriodicBC, RungeKuttaLS, SimpleAverage, StoreSome,
                                      ZeroGradientOutletBC)
    _announce("HIPTarget sod script equations 512 x 60")
    alg = _script_block(1, _SOD_INIT, [ZeroGradientOutletBC(0, 0), ZeroGradientOutletBC(0, 1)],
                        [LFWeno(5, formulation="Z", averaging=SimpleAverage([0, 1])), RungeKuttaLS(3)], True)
    with HIPTarget(alg, parameters=_SOD_PARAMS) as t:
        assert t.lib.path.endswith("libosb_hip.so")
        q = t.initialise().run().state()
    want = run(oracle, P.sod(n=512, steps=60))[0]
    assert np.array_equal(q, want)

    _announce("HIPTarget tgv script equations 24^3 x 6")
    alg = _script_block(3, _TGV_INIT, [PeriodicBC(d, s) for d in range(3) for s in range(2)],
                        [StoreSome(4, "u0 u1 u2"), RungeKuttaLS(3)])
    params = {"gama": "1.4", "Minf": "0.1", "Re": "1600.0", "Pr": "0.71", "dt": "0.018", "niter": "6"}
    for d in range(3):
        params[f"block0np{d}"] = "24"
        params[f"Delta{d}block0"] = f"2*M_PI/block0np{d}"
    NS = NS_Split("Feiereisen", 3, ["Re", "Pr", "gama", "Minf"], viscosity="constant")
    with HIPTarget(alg, parameters=params, physics=NS) as t:
        q0 = initial_fields(alg.blocks[0], t.parameters, t.desc)
        q = t.initialise().run().state()
    ref = P.tgv(n=24, steps=6, dt=0.018)
    assert np.max(np.abs(q0 - ref.q0) / np.maximum(np.abs(ref.q0), 1.0)) < 1e-15
    b, _ = run(oracle, P.Problem("tgv_script", ref.desc, q0, 6))
    l2, li = rel_errors(q, b)
    print(f"   rel L2 {l2.max():.3e} rel Linf {li.max():.3e}", flush=True)
    assert l2.max() <= 1e-11 and li.max() <= 1e-11
