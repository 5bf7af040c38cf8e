"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
oracle and the HDF5 I/O library, each compiled with -fsanitize=address,undefined
into a driver that runs it through its C ABI, and the stage kernels' tile-order
arithmetic (osb_order.hpp) and the eigenvector structural zeros
(osb_eigzero.hpp, against the oracle's eigensystem) in drivers of their own
(tests/sanitize/).  A sanitizer
report, a failed run or a non-finite result fails the test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def _build(out, target):
    r = subprocess.run(["make", "-C", SAN, f"OUT={out}", os.path.join(str(out), target)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stdout[-2000:] + r.stderr[-2000:])


def _run(cmd, tmp):
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=str(tmp), timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "all ok" in r.stdout, out[-4000:]
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    _build(tmp_path, "oracle_driver")
    out = _run([str(tmp_path / "oracle_driver")], tmp_path)
    assert out.count("\nok ") + out.startswith("ok ") >= 13


def test_io_library_under_sanitizers(tmp_path):
    prefix = os.environ.get("HDF5_PREFIX", "/opt/conda")
    if not os.path.exists(os.path.join(prefix, "include", "hdf5.h")):
        pytest.skip("no HDF5 C library")
    _build(tmp_path, "io_driver")
    _run([str(tmp_path / "io_driver"), str(tmp_path)], tmp_path)


def test_tile_order_under_sanitizers(tmp_path):
    """Fast quotients against integer division; every tile once, in range, the
    same as the plain division form, with each mode's XCD locality; the
    kernels' flat index splits against integer division."""
    _build(tmp_path, "order_driver")
    out = _run([str(tmp_path / "order_driver")], tmp_path)
    assert "ok cf_div" in out and "ok tile walks" in out and "ok index splits" in out


def test_eigenvector_structural_zeros_under_sanitizers(tmp_path):
    """The L and R entries that the face kernels skip are exactly the zero
    entries of the oracle's eigensystem, in 1-D, 2-D and 3-D along every
    direction, on random states."""
    _build(tmp_path, "eigzero_driver")
    out = _run([str(tmp_path / "eigzero_driver")], tmp_path)
    assert "ok structural zeros" in out
    assert "ndim 3 dir 2: 6 L and 6 R zero entries per state" in out, out
