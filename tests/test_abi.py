"""CPU tests of the C ABI: both libraries load and export what include/osb.h
declares, the descriptor layout agrees between C and Python, and invalid
descriptors are rejected with the documented error codes (oracle only: the HIP
library is not called without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from opensbli_amd import native as nat, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "osb.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(osb_[a-z0-9_]+)\s*\(", text)))


def declared_macros():
    vals = {}
    for m in re.finditer(r"#define\s+OSB_([A-Z0-9_]+)\s+(-?\d+)", open(HEADER).read()):
        vals[m.group(1)] = int(m.group(2))
    return vals


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_header_declares_the_core_api():
    fns = declared_functions()
    for f in ("osb_ctx_create", "osb_step", "osb_residual", "osb_batch_step", "osb_profile_read"):
        assert f in fns


def test_hip_library_exports_every_declared_function(hip_so):
    missing = [f for f in declared_functions() if f not in exported(hip_so)]
    assert not missing, missing


def test_oracle_exports_the_single_block_api(oracle):
    """The oracle restates one block; timing, batching and the pool are HIP-library only."""
    single = [f for f in declared_functions() if not f.startswith(("osb_batch_", "osb_profile_", "osb_pool_"))]
    missing = [f for f in single if f not in exported(oracle.path)]
    assert not missing, missing


def test_io_library_exports_every_declared_function():
    """libosb_io.so (HDF5 output/restart) exports what include/osb_io.h declares."""
    from opensbli_amd import io_hdf5
    if not os.path.exists(io_hdf5.IO_LIBRARY):
        pytest.skip("libosb_io.so not built (no HDF5 C library)")
    fns = declared_functions(os.path.join(ROOT, "include", "osb_io.h"))
    assert "osb_io_write" in fns and "osb_io_read_dat" in fns
    missing = [f for f in fns if f not in exported(io_hdf5.IO_LIBRARY)]
    assert not missing, missing
    io_hdf5.load_io()


def test_hip_library_loads_without_a_gpu(hip_so):
    lib = C.CDLL(hip_so)
    assert lib.osb_abi_version() == nat.ABI_VERSION


def test_python_constants_match_header():
    for name, val in declared_macros().items():
        if name == "ABI_VERSION":
            assert nat.ABI_VERSION == val
        elif name.startswith("E_") or name == "OK":
            assert getattr(nat, name) == val, name
        else:
            assert getattr(nat, name) == val, name


def test_descriptor_struct_layout():
    """Field offsets of osb_problem_desc as compiled by the C compiler."""
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "osb.h"
#define O(f) printf("%s %zu\n", #f, offsetof(osb_problem_desc, f));
#define S(f) printf("%s %zu\n", #f, offsetof(osb_stats, f));
int main(void) {
  printf("size %zu\n", sizeof(osb_problem_desc));
  O(ndim) O(n) O(delta) O(conv_scheme) O(bc) O(gama) O(weno_eps) O(teno_ct) O(dt) O(cfl) O(rk_stages) O(flux_type) O(rk_a) O(rk_b) O(bc_value) O(shock_filter) O(filter_order) O(filter_sensor) O(ducros_check) O(ducros_select)
  printf("ssize %zu\n", sizeof(osb_stats));
  S(iter) S(time) S(last_dt) S(halo) S(nv) S(device_bytes)
  return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cfile, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(cfile, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cfile, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    c_off = {ln.split()[0]: int(ln.split()[1]) for ln in out.splitlines()}
    assert c_off["size"] == C.sizeof(nat.ProblemDescC)
    assert c_off["ssize"] == C.sizeof(nat.StatsC)
    for f in ("ndim", "n", "delta", "conv_scheme", "bc", "gama", "weno_eps", "teno_ct", "dt", "cfl",
              "rk_stages", "flux_type", "rk_a", "rk_b", "bc_value", "shock_filter", "filter_order", "filter_sensor",
              "ducros_check", "ducros_select"):
        assert getattr(nat.ProblemDescC, f).offset == c_off[f], f
    for f in ("iter", "time", "last_dt", "halo", "nv", "device_bytes"):
        assert getattr(nat.StatsC, f).offset == c_off[f], f


def _bad(desc, **kw):
    d = nat.ProblemDesc(**{**desc.__dict__, **kw})
    return d


@pytest.mark.parametrize("change", [
    dict(ndim=4), dict(n=[2]), dict(delta=[0.0]), dict(conv_scheme=9), dict(order=7),
    dict(bc=[[nat.BC_PERIODIC, nat.BC_ZERO_GRADIENT]]), dict(dt=0.0, cfl=0.0), dict(gama=1.0),
    dict(precision=3), dict(time_scheme=0),
])
def test_invalid_descriptors_rejected(oracle, change):
    base = P.sod(n=32).desc
    d = _bad(base, **change)
    with pytest.raises(nat.OSBError) as e:
        oracle.context(d)
    assert e.value.code in (nat.E_INVALID, nat.E_UNSUPPORTED)


def test_struct_size_checked(oracle):
    d = P.sod(n=32).desc.to_c()
    d.struct_size = 8
    h = C.c_void_p()
    assert oracle.lib.osb_ctx_create(C.byref(d), C.byref(h)) == nat.E_INVALID
    assert oracle.lib.osb_last_error(None)


def test_state_layout_roundtrip(oracle):
    pr = P.tgv(n=12)      # periodic: no boundary point is rewritten by a BC
    with oracle.context(pr.desc) as c:
        c.set_state(pr.q0)
        assert np.array_equal(c.get_state(), pr.q0)
        st = c.stats()
        assert st.nv == 5 and st.iter == 0


def test_nan_check_reports_nonfinite(oracle):
    pr = P.sod(n=32)
    q = pr.q0.copy()
    q[0, 0, 0, 5] = np.nan
    with oracle.context(pr.desc) as c:
        c.set_state(q)
        assert c.nan_check() == nat.E_NONFINITE
    with oracle.context(pr.desc) as c:
        c.set_state(pr.q0)
        assert c.nan_check() == nat.OK


@pytest.mark.parametrize("maker", [
    lambda: P.sod(n=64, steps=3),
    lambda: P.kelvin_helmholtz(n=16, steps=2),
    lambda: P.tgv(n=12, steps=2, scheme="teno", minf=1.25, cfl=0.5, dt=0.0),
])
def test_stages_equal_steps(oracle, maker):
    """osb_stage over every stage of every step equals osb_step bit for bit
    (iteration count and time included); out-of-order stages and a step
    started mid-way are refused."""
    pr = maker()
    nst = 3
    with oracle.context(pr.desc) as a, oracle.context(pr.desc) as b:
        a.set_state(pr.q0)
        b.set_state(pr.q0)
        a.step(pr.steps)
        for _ in range(pr.steps):
            for s in range(nst):
                b.stage(s)
        assert np.array_equal(a.get_state(), b.get_state())
        sa, sb = a.stats(), b.stats()
        assert sa.iter == sb.iter and sa.time == sb.time
        with pytest.raises(nat.OSBError):
            b.stage(1)
        b.stage(0)
        with pytest.raises(nat.OSBError):
            b.step(1)
        b.stage(1)
        b.stage(2)
        b.step(1)


def test_first_rkls_stage_is_b0_dt_residual(oracle):
    """RK-LS stage 0 from a fresh state (du = 0, A0 = 0): q1 = q0 + B0 dt R(q0),
    B0 = 1/3 (rk_LS.py:121-124)."""
    pr = P.kelvin_helmholtz(n=16, steps=1)
    with oracle.context(pr.desc) as c:
        c.set_state(pr.q0)
        R = c.residual()
        c.set_state(pr.q0)
        c.stage(0)
        q1 = c.get_state()
    want = pr.q0 + (1.0 / 3.0) * (pr.desc.dt * R)
    inner = (slice(None), slice(None), slice(1, -1))      # the wall rows are rewritten by the BC after the stage
    assert np.allclose(q1[inner], want[inner], rtol=1e-15, atol=1e-15 * np.abs(pr.q0).max())
    assert not np.array_equal(q1, pr.q0)
