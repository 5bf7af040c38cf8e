"""HDF5 output and restart in the OPS layout (libosb_io.so, opensbli_amd/io_hdf5.py).

The layout is checked against a description of a reference-held OPS block
file (tests/golden/ops_hdf5_layout.json, made by make_ops_hdf5_layout.py from
apps/cylinder/incompressible_cylinder/data.h5) with an independent reader,
h5py under the image's conda Python.  Restart is checked by continuation:
n steps, write, restart, m steps equals n + m steps bit for bit."""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import run
from opensbli_amd import io_hdf5 as io, native as nat, problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
H5PY_PYTHON = "/opt/conda/bin/python3.9"


def _h5py(path, what="layout"):
    if not os.path.exists(H5PY_PYTHON):
        pytest.skip("no Python with h5py in this image")
    out = subprocess.run([H5PY_PYTHON, os.path.join(HERE, "golden", "h5layout.py"), path, what],
                         capture_output=True, text=True, cwd="/tmp")
    if out.returncode != 0:
        pytest.skip(f"h5py not importable: {out.stderr[-200:]}")
    return json.loads(out.stdout)


@pytest.fixture(scope="module")
def io_lib():
    try:
        return io.load_io()
    except FileNotFoundError as e:
        pytest.skip(str(e))


def _written(oracle, tmp_path, pr, steps=3, name="out.h5"):
    path = str(tmp_path / name)
    with oracle.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(steps)
        io.write_state(path, c, constants={"gama": pr.desc.gama, "niter": steps, "dt": pr.desc.dt},
                       int_constants={"niter"})
        q = c.get_state()
        qp = c.get_state(nat.LAYOUT_PADDED)
        st = c.stats()
    return path, q, qp, st


def test_layout_matches_the_reference_file(oracle, io_lib, tmp_path):
    """Same groups/attributes, attribute types, sizes, string padding and axis
    order as the reference-held file; our dataset names are the state's."""
    ref = json.load(open(os.path.join(HERE, "golden", "ops_hdf5_layout.json")))
    pr = P.riemann_2d(n=20, steps=2)
    path, q, qp, st = _written(oracle, tmp_path, pr, 2)
    got = _h5py(path)
    assert got["opensbliblock00"] == ref["layout"]["opensbliblock00"]
    want_ds = ref["layout"]["opensbliblock00/x0_B0"]
    for nm in io.conservative_names(2):
        assert got[f"opensbliblock00/{nm}"] == want_ds, nm
    vals = _h5py(path, "values")
    a = vals["attrs"]["opensbliblock00/rho_B0"]
    r = ref["attr_values"]["opensbliblock00/x0_B0"]
    assert {k: a[k] for k in ("base", "block", "block_index", "dim", "ops_type", "type")} == \
           {k: r[k] for k in ("base", "block", "block_index", "dim", "ops_type", "type")}
    # axis order: the reference's shape is slowest axis first, size x first
    assert ref["x0_B0_shape"][::-1] == [s - m + p for s, m, p in zip(r["size"], r["d_m"], r["d_p"])]
    h = int(st.halo)
    assert a["size"] == [20, 20] and a["d_m"] == [-h, -h] and a["d_p"] == [h, h]
    assert vals["consts"]["iter"] == {"value": 2, "dtype": "<i4"}
    assert vals["consts"]["niter"]["dtype"] == "<i4" and vals["consts"]["gama"] == {"value": 1.4, "dtype": "<f8"}
    assert vals["consts"]["simulation_time"]["value"] == st.time


def test_roundtrip_bitwise(oracle, io_lib, tmp_path):
    pr = P.kelvin_helmholtz(n=24, steps=2)
    path, q, qp, st = _written(oracle, tmp_path, pr, 2)
    q2, it, t = io.read_state(path, pr.desc)
    assert np.array_equal(q, q2) and it == 2 and t == st.time
    for v, nm in enumerate(io.conservative_names(2)):
        a, size, dm, dp = io.read_dat(path, nm)
        assert np.array_equal(a, qp[v, 0]) and size == [24, 24]
    assert io.read_const(path, "gama") == 1.4 and io.read_const(path, "niter") == 2.0


def test_reads_a_file_written_by_h5py(io_lib, tmp_path):
    """A block file written independently (h5py, with the reference's metadata
    set and a halo depth of 5 as in its data files) reads back with the right
    interior."""
    if not os.path.exists(H5PY_PYTHON):
        pytest.skip("no Python with h5py in this image")
    path = str(tmp_path / "ext.h5")
    script = f"""
import h5py, numpy as np
n0, n1, h = 12, 9, 5
a = np.arange((n1 + 2*h)*(n0 + 2*h), dtype=np.float64).reshape(n1 + 2*h, n0 + 2*h)
with h5py.File({path!r}, 'w') as f:
    g = f.create_group('opensbliblock00')
    g.attrs.create('dims', [2], dtype='int32'); g.attrs.create('ops_type', 'ops_block', dtype='S9')
    g.attrs.create('index', [0], dtype='int32')
    for nm, k in (('rho_B0', 1.0), ('rhou0_B0', 2.0), ('rhou1_B0', 3.0), ('rhoE_B0', 4.0)):
        d = g.create_dataset(nm, data=k * a)
        d.attrs.create('d_p', [h, h], dtype='int32'); d.attrs.create('d_m', [-h, -h], dtype='int32')
        d.attrs.create('dim', [1], dtype='int32'); d.attrs.create('ops_type', 'ops_dat', dtype='S10')
        d.attrs.create('block_index', [0], dtype='int32'); d.attrs.create('base', [0, 0], dtype='int32')
        d.attrs.create('type', 'double', dtype='S15'); d.attrs.create('block', 'opensbliblock00', dtype='S25')
        d.attrs.create('size', [n0, n1], dtype='int32')
    f.create_dataset('iter', data=np.array([7], dtype='int32'))
    f.create_dataset('simulation_time', data=np.array([0.25]))
"""
    r = subprocess.run([H5PY_PYTHON, "-c", script], capture_output=True, text=True, cwd="/tmp")
    if r.returncode != 0:
        pytest.skip(f"h5py not usable: {r.stderr[-200:]}")
    desc = nat.ProblemDesc(ndim=2, n=[12, 9], delta=[0.1, 0.1], conv_scheme=nat.CONV_WENO_Z, dt=1e-3)
    q, it, t = io.read_state(path, desc)
    a = np.arange(19 * 22, dtype=np.float64).reshape(19, 22)[5:14, 5:17]
    for v in range(4):
        assert np.array_equal(q[v, 0], (v + 1.0) * a)
    assert it == 7 and t == 0.25


@pytest.mark.parametrize("maker", [lambda: P.sod(n=64, steps=20), lambda: P.kelvin_helmholtz(n=24, steps=12),
                                   lambda: P.tgv(n=12, steps=6)])
def test_restart_continues_bit_for_bit(oracle, io_lib, tmp_path, maker):
    pr = maker()
    k = pr.steps // 2
    straight, st = run(oracle, pr, pr.steps)
    path = str(tmp_path / "restart.h5")
    with oracle.context(pr.desc) as c:
        c.set_state(pr.q0)
        c.step(k)
        io.write_state(path, c)
    with oracle.context(pr.desc) as c:
        it, t = io.restart(c, path)
        assert it == k
        c.step(pr.steps - k)
        q = c.get_state()
        s2 = c.stats()
    assert np.array_equal(q, straight)
    assert s2.iter == st.iter and s2.time == st.time


def test_target_writes_where_the_time_loop_does(oracle, io_lib, tmp_path):
    """iohdf5(save_every=4) over 10 iterations: files after loop indices 0, 3
    and 7 (named 1, 4, 8) carrying iter = the loop index, and the final file
    with iter = 10 (algorithm.py:476-485, io_hdf5.py:190-206); a second target
    restarted from the final file continues bit for bit."""
    from opensbli_amd.code_generation.hip import HIPTarget
    from opensbli_amd.surface import (LFWeno, RungeKuttaLS, SimpleAverage, SimulationBlock,
                                      TraditionalAlgorithmRK, ZeroGradientOutletBC, iohdf5)

    def alg(ios):
        block = SimulationBlock(1, block_number=0)
        block.set_block_boundaries([ZeroGradientOutletBC(0, 0), ZeroGradientOutletBC(0, 1)])
        weno = LFWeno(5, formulation="Z", averaging=SimpleAverage([0, 1]))
        block.set_discretisation_schemes({"weno": weno, "rk": RungeKuttaLS(3)})
        block.setio(ios)
        return TraditionalAlgorithmRK(block)

    params = {"gama": "1.4", "dt": "0.0002", "niter": "10", "block0np0": "128", "Delta0block0": "1.0/(block0np0-1)"}
    pr = P.sod(n=128)
    out = tmp_path / "run"
    out.mkdir()
    t = HIPTarget(alg([iohdf5(save_every=4, iotype="Write")]), parameters=params, library=oracle, output_dir=str(out))
    t.initialise(pr.q0).run()
    t.close()
    names = sorted(os.listdir(out))
    assert names == ["opensbli_output.h5", "opensbli_output_000001.h5", "opensbli_output_000004.h5",
                     "opensbli_output_000008.h5"], names
    for nm, it in (("opensbli_output_000001.h5", 0), ("opensbli_output_000004.h5", 3),
                   ("opensbli_output_000008.h5", 7), ("opensbli_output.h5", 10)):
        assert io.read_const(str(out / nm), "iter") == it, nm
    assert io.read_const(str(out / "opensbli_output.h5"), "block0np0") == 128
    _, st = run(oracle, pr, 10)
    assert io.read_const(str(out / "opensbli_output.h5"), "simulation_time") == st.time
    # restart from the final file: iohdf5(iotype='Read', restart_simulation=True) reads restart.h5
    os.rename(out / "opensbli_output.h5", out / "restart.h5")
    t2 = HIPTarget(alg([iohdf5(iotype="Read", restart_simulation="True")]), parameters=params, library=oracle,
                   output_dir=str(out))
    q = t2.run().state()
    assert t2.stats().iter == 20
    t2.close()
    straight, _ = run(oracle, pr, 20)
    assert np.array_equal(q, straight)
