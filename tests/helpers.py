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


def fused_3d(desc):
    """Whether the HIP library steps `desc` through its fused 3-D stage kernel
    (3-D central schemes, and 3-D shock capturing with viscous terms, away from
    Carpenter closure faces).  That kernel evaluates the constituent relations
    with one reciprocal of rho per point and central weights pre-scaled by
    1/Delta, so it meets the oracle at the FP64 bar (relative L2 and L-inf
    <= 1e-11) instead of bit for bit; every other path rounds like the oracle."""
    from opensbli_amd import native as nat
    central = desc.conv_scheme in (nat.CONV_CENTRAL_FEIEREISEN, nat.CONV_CENTRAL_SKEW, nat.CONV_CENTRAL_KGP)
    closure = any(desc.bc[d][s] >= nat.BC_EXTRAPOLATION0 for d in range(desc.ndim) for s in range(2))
    return desc.ndim == 3 and (central or desc.viscous != nat.VISC_NONE) and not closure


def roundoff_spread(lib, problem, steps=None, seeds=(1, 2)):
    """Per-variable relative L2 and L-inf by which `lib`'s own result moves when
    the initial state is perturbed by about one part in 1e16 (seeded): the
    roundoff conditioning of the problem over those steps.  Threshold
    selections (the Ducros hybrid filter, the TENO cut-off) and variables that
    are analytically zero make it large; smooth runs keep it near 1e-15."""
    base, _ = run(lib, problem, steps)
    l2s, lis = [], []
    for s in seeds:
        rng = np.random.default_rng(s)
        b, _ = run(lib, problem, steps, state=problem.q0 * (1.0 + 1e-16 * rng.standard_normal(problem.q0.shape)))
        l2, li = rel_errors(b, base)
        l2s.append(l2)
        lis.append(li)
    return np.max(l2s, axis=0), np.max(lis, axis=0)


def fp64_bar(a, b, desc, oracle=None, problem=None, steps=None, tol=1e-11):
    """The FP64 parity bar: relative L2 and L-inf <= tol per variable.  On the
    fused 3-D path (which does not round like the oracle) a variable that
    misses tol is held to 10x the oracle's own roundoff spread instead
    (roundoff_spread), when that is larger: the two libraries then differ by
    less than the problem's response to one-ulp changes of its input.
    Returns the largest relative bound applied (tol when the case met it)."""
    l2, li = rel_errors(a, b)
    if l2.max() <= tol and li.max() <= tol:
        return tol
    assert fused_3d(desc) and oracle is not None, (l2, li)
    s2, si = roundoff_spread(oracle, problem, steps)
    b2, bi = np.maximum(tol, 10.0 * s2), np.maximum(tol, 10.0 * si)
    fmt = {"float_kind": lambda v: f"{v:.2e}"}
    print(f"   roundoff-conditioned: oracle spread L2 {s2.max():.3e} Linf {si.max():.3e}; "
          f"bound per variable L2 {np.array2string(b2, formatter=fmt)} Linf {np.array2string(bi, formatter=fmt)}", flush=True)
    assert np.all(l2 <= b2) and np.all(li <= bi), (l2, li, b2, bi)
    return float(bi.max())


def assert_exact_unless_fused(a, b, desc, tol=1e-11):
    """Bitwise equality with the oracle, or the FP64 bar on the fused 3-D path."""
    if fused_3d(desc):
        l2, li = rel_errors(a, b)
        assert l2.max() <= tol and li.max() <= tol, (l2, li)
    else:
        assert np.array_equal(a, b), float(np.abs(a - b).max())


def kept_fastest(choice, options, times_us, resolution=0.01):
    """An autotuner's choice among `options` is one whose reported time (printed
    to `resolution` microseconds) is the smallest, up to that rounding: two
    candidates whose times round alike may be kept either way."""
    assert choice in options, (choice, options)
    t = times_us[options.index(choice)]
    return t <= min(times_us) + resolution + 1e-9
