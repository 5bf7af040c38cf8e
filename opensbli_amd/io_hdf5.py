"""HDF5 output and restart in the OPS layout the generated OpenSBLI program uses.

The files are what ``ops_fetch_block_hdf5_file`` / ``ops_fetch_dat_hdf5_file``
/ ``write_constants`` write and ``ops_decl_dat_hdf5`` / ``ops_get_const_hdf5``
read back on restart (opensbli/core/io_hdf5.py:165-250,
opensbli/code_generation/opsc.py:962-977, :1198-1252): a group
``opensbliblock00`` holding ``rho_B0``, ``rhou0_B0``, ..., ``rhoE_B0`` as full
padded arrays (slowest axis first) with the OPS metadata attributes, and the
simulation constants as one-element datasets at the root.  The writing and
reading are done by ``libosb_io.so`` (include/osb_io.h).

Iteration bookkeeping follows the generated time loop
(opensbli/code_generation/algorithm/algorithm.py:180-185, :476-485): the
final file carries ``iter`` = the completed iterations and the simulation
time; a file written inside the loop after iteration k (1-based, named
``opensbli_output_%06d.h5`` % k) carries ``iter`` = k - 1, the loop index at
that point.  A restart reads ``iter`` into the start iteration and
``simulation_time`` into the start time.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Iterable, Mapping, Optional, Sequence, Tuple

import numpy as np

from . import native as nat

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIBRARY = os.path.join(_HERE, "libosb_io.so")


class IOBlockC(C.Structure):
    _fields_ = [("name", C.c_char_p), ("index", C.c_int32), ("ndim", C.c_int32),
                ("size", C.c_int32 * 3), ("halo", C.c_int32)]


_lib = None


def load_io(path: Optional[str] = None):
    """The HDF5 library binding; raises if libosb_io.so was not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or IO_LIBRARY
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not found: build it (needs the HDF5 C library, HDF5_PREFIX in csrc/Makefile)")
    L = C.CDLL(path)
    cpp = C.POINTER(C.c_char_p)
    dpp = C.POINTER(C.POINTER(C.c_double))
    i32 = C.POINTER(C.c_int32)
    L.osb_io_write.argtypes = [C.c_char_p, C.POINTER(IOBlockC), C.c_int, cpp, dpp, C.c_int, cpp,
                               C.POINTER(C.c_double), C.c_int, cpp, i32]
    L.osb_io_dat_info.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, i32, i32, i32, i32]
    L.osb_io_read_dat.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_int64]
    L.osb_io_read_const.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double)]
    L.osb_io_last_error.restype = C.c_char_p
    if path == IO_LIBRARY:
        _lib = L
    return L


def _check(L, rc):
    if rc != 0:
        raise nat.OSBError(rc, L.osb_io_last_error().decode())


def conservative_names(ndim: int, block_number: int = 0):
    """Dataset names of the time-advanced arrays, rho_B0, rhou0_B0, ..., rhoE_B0."""
    return [f"{v}_B{block_number}" for v in ["rho"] + [f"rhou{d}" for d in range(ndim)] + ["rhoE"]]


def write_block(path: str, arrays: Mapping[str, np.ndarray], ndim: int, size: Sequence[int], halo: int,
                constants: Optional[Mapping[str, float]] = None, int_constants: Iterable[str] = (),
                block_name: str = "opensbliblock00", block_index: int = 0):
    """Write one block's padded arrays ``{name: array}`` (each of shape
    (P2, P1, P0) or flat, P = size + 2 halo) and the root constants; names in
    ``int_constants`` are stored as int32."""
    L = load_io()
    blk = IOBlockC(block_name.encode(), int(block_index), int(ndim), (C.c_int32 * 3)(*(list(size) + [1] * (3 - len(size)))),
                   int(halo))
    P = [int(size[d]) + 2 * int(halo) for d in range(ndim)]
    keep = []
    names = (C.c_char_p * max(len(arrays), 1))()
    ptrs = (C.POINTER(C.c_double) * max(len(arrays), 1))()
    for i, (nm, a) in enumerate(arrays.items()):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != int(np.prod(P)):
            raise ValueError(f"{nm}: {a.size} values, expected {int(np.prod(P))} (padded {P[::-1]})")
        keep.append(a)
        names[i] = nm.encode()
        ptrs[i] = a.ctypes.data_as(C.POINTER(C.c_double))
    constants = dict(constants or {})
    ints = set(int_constants)
    dn = [k for k in constants if k not in ints]
    inn = [k for k in constants if k in ints]
    dnames = (C.c_char_p * max(len(dn), 1))(*[k.encode() for k in dn])
    dvals = (C.c_double * max(len(dn), 1))(*[float(constants[k]) for k in dn])
    inames = (C.c_char_p * max(len(inn), 1))(*[k.encode() for k in inn])
    ivals = (C.c_int32 * max(len(inn), 1))(*[int(constants[k]) for k in inn])
    _check(L, L.osb_io_write(os.fsencode(path), C.byref(blk), len(arrays), names, ptrs, len(dn), dnames, dvals,
                             len(inn), inames, ivals))


def read_dat(path: str, name: str, block_name: str = "opensbliblock00") -> Tuple[np.ndarray, list, list, list]:
    """(padded array shaped slowest-axis first, size, d_m, d_p) of one dataset."""
    L = load_io()
    nd = C.c_int32()
    size, dm, dp = (C.c_int32 * 3)(), (C.c_int32 * 3)(), (C.c_int32 * 3)()
    _check(L, L.osb_io_dat_info(os.fsencode(path), block_name.encode(), name.encode(), C.byref(nd), size, dm, dp))
    n = nd.value
    shape = [size[d] - dm[d] + dp[d] for d in range(n)][::-1]
    out = np.empty(shape, dtype=np.float64)
    _check(L, L.osb_io_read_dat(os.fsencode(path), block_name.encode(), name.encode(),
                                out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
    return out, list(size)[:n], list(dm)[:n], list(dp)[:n]


def read_const(path: str, name: str) -> float:
    L = load_io()
    v = C.c_double()
    _check(L, L.osb_io_read_const(os.fsencode(path), name.encode(), C.byref(v)))
    return v.value


def write_state(path: str, ctx: "nat.Context", constants: Optional[Mapping[str, float]] = None,
                int_constants: Iterable[str] = (), iteration: Optional[int] = None, block_number: int = 0):
    """Write a context's conservative variables (padded, at the library's halo
    depth) with the simulation constants, ``iter`` (int) and
    ``simulation_time`` -- what the generated program writes at the end of a
    run.  ``iteration`` overrides the ``iter`` value (the in-loop files)."""
    desc = ctx.desc
    st = ctx.stats()
    q = ctx.get_state(nat.LAYOUT_PADDED)
    arrays = {nm: q[v] for v, nm in enumerate(conservative_names(desc.ndim, block_number))}
    consts = dict(constants or {})
    consts["iter"] = int(st.iter if iteration is None else iteration)
    consts["simulation_time"] = float(st.time)
    write_block(path, arrays, desc.ndim, list(desc.n)[:desc.ndim], int(st.halo), consts,
                set(int_constants) | {"iter"}, block_name=f"opensbliblock{block_number:02d}", block_index=block_number)


def read_state(path: str, desc: "nat.ProblemDesc", block_number: int = 0) -> Tuple[np.ndarray, int, float]:
    """Interior state (nv, n2, n1, n0), start iteration and start time from an
    output/restart file (any halo depth)."""
    nv = desc.nv
    n = list(desc.n)[:desc.ndim]
    q = np.empty(desc.shape, dtype=np.float64)
    for v, nm in enumerate(conservative_names(desc.ndim, block_number)):
        a, size, dm, dp = read_dat(path, nm, f"opensbliblock{block_number:02d}")
        if size != n:
            raise ValueError(f"{path}:{nm} has size {size}, the block has {n}")
        sl = tuple(slice(-dm[d], -dm[d] + n[d]) for d in range(desc.ndim))[::-1]
        q[v] = a[sl].reshape(q.shape[1:])
    it = int(read_const(path, "iter"))
    t = read_const(path, "simulation_time")
    return q, it, t


def restart(ctx: "nat.Context", path: str = "restart.h5", block_number: int = 0):
    """Initialise a context from a restart file (opsc.py:962-977, :1198-1252):
    the state, the start iteration and the simulation time."""
    q, it, t = read_state(path, ctx.desc, block_number)
    ctx.set_state(q)
    ctx.set_time(t, it)
    return it, t
