"""ctypes binding of the C ABI declared in include/osb.h.

The same binding drives two libraries with identical entry points:

* ``libosb_hip.so``  -- the product (HIP kernels for gfx950), located next to
  this file and built by ``opensbli_amd/csrc/Makefile``;
* the CPU oracle library, which only the tests and the benchmark's CPU
  baseline load by explicit path.

Nothing here falls back from one to the other: ``load_hip()`` raises if the
HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

ABI_VERSION = 5

OK = 0
E_INVALID = -1
E_UNSUPPORTED = -2
E_DEVICE = -3
E_NOMEM = -4
E_NONFINITE = -5

CONV_WENO_Z = 1
CONV_TENO = 2
CONV_CENTRAL_FEIEREISEN = 3
CONV_CENTRAL_SKEW = 4
CONV_CENTRAL_KGP = 5

AVG_SIMPLE = 1
AVG_ROE = 2

VISC_NONE = 0
VISC_CONSTANT = 1
VISC_SUTHERLAND = 2

TIME_RKLS3 = 1
TIME_RK3 = 2

BC_PERIODIC = 1
BC_ZERO_GRADIENT = 2
BC_INVISCID_WALL = 3
BC_SYMMETRY = 4
BC_EXTRAPOLATION0 = 5
BC_EXTRAPOLATION1 = 6
BC_DIRICHLET = 7

FLUX_LLF = 0
FLUX_GLF = 1
FLUX_HLLC = 2
FLUX_HLLC_LM = 3

FILTER_NONE = 0
FILTER_WENOZ = 1

SENSOR_NONE = 0
SENSOR_DUCROS = 1

PREC_FP64 = 1
PREC_MIXED = 2

LAYOUT_INTERIOR = 0
LAYOUT_PADDED = 1


class ProblemDescC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("abi_version", C.c_uint32),
        ("ndim", C.c_int32),
        ("n", C.c_int32 * 3),
        ("delta", C.c_double * 3),
        ("conv_scheme", C.c_int32),
        ("order", C.c_int32),
        ("averaging", C.c_int32),
        ("viscous", C.c_int32),
        ("time_scheme", C.c_int32),
        ("precision", C.c_int32),
        ("bc", (C.c_int32 * 2) * 3),
        ("gama", C.c_double),
        ("minf", C.c_double),
        ("re", C.c_double),
        ("pr", C.c_double),
        ("weno_eps", C.c_double),
        ("teno_eps", C.c_double),
        ("teno_ct", C.c_double),
        ("shock_filter_control", C.c_double),
        ("dt", C.c_double),
        ("cfl", C.c_double),
        ("rk_stages", C.c_int32),
        ("flux_type", C.c_int32),
        ("rk_a", C.c_double * 5),
        ("rk_b", C.c_double * 5),
        ("bc_value", ((C.c_double * 5) * 2) * 3),
        ("shock_filter", C.c_int32),
        ("filter_order", C.c_int32),
        ("filter_sensor", C.c_int32),
        ("reserved0", C.c_int32),
        ("ducros_check", C.c_double),
        ("ducros_select", C.c_double),
    ]


class StatsC(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("iter", C.c_int64),
        ("time", C.c_double),
        ("last_dt", C.c_double),
        ("halo", C.c_int32),
        ("nv", C.c_int32),
        ("device_bytes", C.c_int64),
    ]


@dataclass
class ProblemDesc:
    """Python mirror of ``osb_problem_desc``."""

    ndim: int
    n: Sequence[int]
    delta: Sequence[float]
    conv_scheme: int
    order: int = 5
    averaging: int = AVG_SIMPLE
    viscous: int = VISC_NONE
    time_scheme: int = TIME_RKLS3
    precision: int = PREC_FP64
    bc: Sequence[Sequence[int]] = field(default_factory=lambda: [[BC_PERIODIC, BC_PERIODIC]] * 3)
    gama: float = 1.4
    minf: float = 1.0
    re: float = 1.0
    pr: float = 0.71
    weno_eps: float = 1.0e-40
    teno_eps: float = 1.0e-40
    teno_ct: float = 1.0e-5
    shock_filter_control: float = 1.0
    dt: float = 0.0
    cfl: float = 0.0
    # RungeKuttaLS coefficients; empty = the 3-stage Williamson scheme
    rk_a: Sequence[float] = ()
    rk_b: Sequence[float] = ()
    flux_type: int = FLUX_LLF
    shock_filter: int = FILTER_NONE
    # WENO filter: order 5 or 3; SENSOR_DUCROS = WENOFilter(optimize=True)
    filter_order: int = 5
    filter_sensor: int = SENSOR_NONE
    ducros_check: float = 0.05
    ducros_select: float = 0.05
    # BC_DIRICHLET faces: {(direction, side): conservative values (rho, rhou.., rhoE)}
    bc_value: dict = field(default_factory=dict)

    @property
    def nv(self) -> int:
        return self.ndim + 2

    @property
    def shape(self):
        """Interior array shape (nv, n2, n1, n0) -- x fastest."""
        n = list(self.n) + [1] * (3 - len(self.n))
        return (self.nv, n[2], n[1], n[0])

    @property
    def npoints(self) -> int:
        n = list(self.n) + [1] * (3 - len(self.n))
        return int(n[0]) * int(n[1]) * int(n[2])

    def to_c(self) -> ProblemDescC:
        d = ProblemDescC()
        d.struct_size = C.sizeof(ProblemDescC)
        d.abi_version = ABI_VERSION
        d.ndim = self.ndim
        n = list(self.n) + [1] * (3 - len(self.n))
        dl = list(self.delta) + [1.0] * (3 - len(self.delta))
        for i in range(3):
            d.n[i] = int(n[i])
            d.delta[i] = float(dl[i])
        d.conv_scheme = self.conv_scheme
        d.order = self.order
        d.averaging = self.averaging
        d.viscous = self.viscous
        d.time_scheme = self.time_scheme
        d.precision = self.precision
        bc = list(self.bc) + [[BC_PERIODIC, BC_PERIODIC]] * (3 - len(self.bc))
        for i in range(3):
            for s in range(2):
                d.bc[i][s] = int(bc[i][s])
        d.gama = self.gama
        d.minf = self.minf
        d.re = self.re
        d.pr = self.pr
        d.weno_eps = self.weno_eps
        d.teno_eps = self.teno_eps
        d.teno_ct = self.teno_ct
        d.shock_filter_control = self.shock_filter_control
        d.dt = self.dt
        d.cfl = self.cfl
        if len(self.rk_a) != len(self.rk_b) or len(self.rk_a) > 5:
            raise ValueError("rk_a and rk_b need the same length, at most 5")
        d.rk_stages = len(self.rk_a)
        for s, (a, b) in enumerate(zip(self.rk_a, self.rk_b)):
            d.rk_a[s] = float(a)
            d.rk_b[s] = float(b)
        d.flux_type = int(self.flux_type)
        d.shock_filter = int(self.shock_filter)
        d.filter_order = int(self.filter_order)
        d.filter_sensor = int(self.filter_sensor)
        d.ducros_check = float(self.ducros_check)
        d.ducros_select = float(self.ducros_select)
        for (i, s), vals in self.bc_value.items():
            vals = [float(v) for v in vals]
            if len(vals) != self.nv:
                raise ValueError(f"bc_value[{i},{s}] needs {self.nv} conservative values")
            for v, x in enumerate(vals):
                d.bc_value[i][s][v] = x
        return d


class OSBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"osb error {code}: {msg}")
        self.code = code


_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIBRARY = os.path.join(_HERE, "libosb_hip.so")


class Library:
    """One loaded implementation of the osb C ABI."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} not found; build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        self.path = path
        self.lib = C.CDLL(path)
        L = self.lib
        vp = C.c_void_p
        L.osb_abi_version.restype = C.c_int
        L.osb_ctx_create.argtypes = [C.POINTER(ProblemDescC), C.POINTER(vp)]
        L.osb_ctx_destroy.argtypes = [vp]
        L.osb_last_error.argtypes = [vp]
        L.osb_last_error.restype = C.c_char_p
        dptr = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.osb_set_state.argtypes = [vp, dptr, C.c_int]
        L.osb_get_state.argtypes = [vp, dptr, C.c_int]
        L.osb_set_time.argtypes = [vp, C.c_double, C.c_int64]
        L.osb_step.argtypes = [vp, C.c_int]
        L.osb_stage.argtypes = [vp, C.c_int]
        L.osb_apply_bcs.argtypes = [vp]
        L.osb_residual.argtypes = [vp, dptr]
        L.osb_max_wavespeed.argtypes = [vp, C.POINTER(C.c_double)]
        L.osb_sync.argtypes = [vp]
        L.osb_get_stats.argtypes = [vp, C.POINTER(StatsC)]
        L.osb_tgv_diagnostics.argtypes = [vp, C.POINTER(C.c_double)]
        L.osb_nan_check.argtypes = [vp]
        L.osb_probe.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.osb_state_device_ptr.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.osb_set_stream.argtypes = [vp, vp]
        if hasattr(L, "osb_set_state_device"):
            L.osb_set_state_device.argtypes = [vp, vp, C.c_int]
            L.osb_get_state_device.argtypes = [vp, vp, C.c_int]
        self.has_profile = hasattr(L, "osb_profile_enable")
        if self.has_profile:
            L.osb_profile_enable.argtypes = [vp, C.c_int]
            L.osb_profile_read.argtypes = [vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        self.has_batch = hasattr(L, "osb_batch_create")
        if self.has_batch:
            L.osb_batch_create.argtypes = [C.POINTER(ProblemDescC), C.c_int, C.POINTER(vp)]
            L.osb_batch_destroy.argtypes = [vp]
            L.osb_batch_set_state.argtypes = [vp, C.c_int, dptr, C.c_int]
            L.osb_batch_get_state.argtypes = [vp, C.c_int, dptr, C.c_int]
            L.osb_batch_step.argtypes = [vp, C.c_int]
            L.osb_batch_sync.argtypes = [vp]
            L.osb_batch_last_error.argtypes = [vp]
            L.osb_batch_last_error.restype = C.c_char_p
            L.osb_batch_get_stats.argtypes = [vp, C.c_int, C.POINTER(StatsC)]
            L.osb_batch_autotune.argtypes = [vp]
            L.osb_batch_describe.argtypes = [vp, C.c_char_p, C.c_int]
        if hasattr(L, "osb_describe"):
            L.osb_describe.argtypes = [vp, C.c_char_p, C.c_int]
        self.has_noalloc = hasattr(L, "osb_step_noalloc")
        if self.has_noalloc:
            L.osb_step_noalloc.argtypes = [C.POINTER(ProblemDescC), vp, vp, C.c_int, vp]
        L.osb_estimate_bytes.argtypes = [C.POINTER(ProblemDescC), C.POINTER(C.c_int64)]
        L.osb_estimate_bytes.restype = C.c_int
        self.has_pool = hasattr(L, "osb_pool_create")
        if self.has_pool:
            L.osb_pool_create.argtypes = [C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]
            L.osb_pool_destroy.argtypes = [vp]
            L.osb_pool_admit.argtypes = [vp, C.POINTER(ProblemDescC), dptr, C.c_int, C.POINTER(C.c_int)]
            L.osb_pool_advance.argtypes = [vp, C.c_int]
            L.osb_pool_poll.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
            L.osb_pool_fetch.argtypes = [vp, C.c_int, dptr, C.POINTER(StatsC)]
            L.osb_pool_release.argtypes = [vp, C.c_int]
            L.osb_pool_usage.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int)]
            L.osb_pool_last_error.argtypes = [vp]
            L.osb_pool_last_error.restype = C.c_char_p
        if L.osb_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{path}: ABI version {L.osb_abi_version()} != {ABI_VERSION}")

    def context(self, desc: ProblemDesc) -> "Context":
        return Context(self, desc)

    def batch(self, descs: Sequence[ProblemDesc]) -> "Batch":
        return Batch(self, descs)

    def pool(self, hbm_budget_bytes: int, max_solves: int = 64, nstreams: int = 4) -> "Pool":
        return Pool(self, hbm_budget_bytes, max_solves, nstreams)

    def step_noalloc(self, desc: ProblemDesc, q_ptr: int, du_ptr: int, nsteps: int, stream_ptr: int = 0):
        """osb_step_noalloc: advance a caller-owned padded state (device
        pointer q_ptr) and RK register (du_ptr) by nsteps on the given HIP
        stream without any device allocation (see include/osb.h)."""
        cd = desc.to_c()
        rc = self.lib.osb_step_noalloc(C.byref(cd), C.c_void_p(q_ptr), C.c_void_p(du_ptr), int(nsteps), C.c_void_p(stream_ptr))
        if rc != OK:
            msg = self.lib.osb_last_error(None)
            raise OSBError(rc, msg.decode() if msg else "")

    def estimate_bytes(self, desc: ProblemDesc) -> int:
        cd = desc.to_c()
        out = C.c_int64()
        rc = self.lib.osb_estimate_bytes(C.byref(cd), C.byref(out))
        if rc != OK:
            msg = self.lib.osb_last_error(None)
            raise OSBError(rc, msg.decode() if msg else "")
        return out.value


_hip_lib: Optional[Library] = None


def load_hip() -> Library:
    """Load the HIP product library (OSB_HIP_LIBRARY overrides the in-tree
    path, e.g. for a differently tuned build).  Raises if it has not been built."""
    global _hip_lib
    if _hip_lib is None:
        # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64
        # with the same soname, and torch cannot initialise on a runtime that
        # was loaded before it.  Load torch's first (when installed); the
        # library then binds to it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _hip_lib = Library(os.environ.get("OSB_HIP_LIBRARY") or HIP_LIBRARY)
    return _hip_lib


class Context:
    """An osb_ctx: one block's resident state on one device (or on the CPU for the oracle)."""

    def __init__(self, lib: Library, desc: ProblemDesc):
        self.lib = lib
        self.desc = desc
        self._cdesc = desc.to_c()
        h = C.c_void_p()
        rc = lib.lib.osb_ctx_create(C.byref(self._cdesc), C.byref(h))
        if rc != OK:
            msg = lib.lib.osb_last_error(None)
            raise OSBError(rc, msg.decode() if msg else "")
        self.h = h

    # context manager ---------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.lib.osb_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != OK:
            msg = self.lib.lib.osb_last_error(self.h)
            raise OSBError(rc, msg.decode() if msg else "")

    # state ---------------------------------------------------------------
    def padded_shape(self):
        """(nv, P2, P1, P0) of the padded layout: n + 2h in active directions,
        h = the library's stored halo depth (``stats().halo``)."""
        h = int(self.stats().halo)
        n = list(self.desc.n) + [1] * (3 - len(self.desc.n))
        P = [n[d] + 2 * h if d < self.desc.ndim else 1 for d in range(3)]
        return (self.desc.nv, P[2], P[1], P[0])

    def set_state(self, q: np.ndarray, layout: int = LAYOUT_INTERIOR):
        q = np.ascontiguousarray(q, dtype=np.float64)
        want = int(np.prod(self.padded_shape())) if layout == LAYOUT_PADDED else self.desc.nv * self.desc.npoints
        if q.size != want:
            raise ValueError(f"state has {q.size} values, expected {want}")
        self._check(self.lib.lib.osb_set_state(self.h, q.reshape(-1), int(layout)))

    def get_state(self, layout: int = LAYOUT_INTERIOR) -> np.ndarray:
        shape = self.padded_shape() if layout == LAYOUT_PADDED else self.desc.shape
        out = np.empty(shape, dtype=np.float64)
        self._check(self.lib.lib.osb_get_state(self.h, out.reshape(-1), int(layout)))
        return out

    def set_time(self, t: float, it: int):
        self._check(self.lib.lib.osb_set_time(self.h, float(t), int(it)))

    def step(self, nsteps: int = 1):
        self._check(self.lib.lib.osb_step(self.h, int(nsteps)))

    def stage(self, s: int):
        """One RK stage of the current step (osb_stage), in order."""
        self._check(self.lib.lib.osb_stage(self.h, int(s)))

    def apply_bcs(self):
        self._check(self.lib.lib.osb_apply_bcs(self.h))

    def residual(self) -> np.ndarray:
        out = np.empty(self.desc.shape, dtype=np.float64)
        self._check(self.lib.lib.osb_residual(self.h, out.reshape(-1)))
        return out

    def max_wavespeed(self) -> float:
        v = C.c_double()
        self._check(self.lib.lib.osb_max_wavespeed(self.h, C.byref(v)))
        return v.value

    def sync(self):
        self._check(self.lib.lib.osb_sync(self.h))

    def stats(self) -> StatsC:
        s = StatsC()
        s.struct_size = C.sizeof(StatsC)
        self._check(self.lib.lib.osb_get_stats(self.h, C.byref(s)))
        return s

    def tgv_diagnostics(self) -> np.ndarray:
        out = (C.c_double * 4)()
        self._check(self.lib.lib.osb_tgv_diagnostics(self.h, out))
        return np.array(list(out))

    def nan_check(self) -> int:
        return self.lib.lib.osb_nan_check(self.h)

    def probe(self, points: Sequence[Sequence[int]], variables: Sequence[int]) -> np.ndarray:
        """Values of conservative variables[p] at interior grid points[p] (i, j, k)."""
        n = len(points)
        if len(variables) != n:
            raise ValueError("one variable per probe point")
        ijk = (C.c_int32 * max(3 * n, 1))()
        for p, pt in enumerate(points):
            pt = list(pt) + [0] * (3 - len(pt))
            for d in range(3):
                ijk[3 * p + d] = int(pt[d])
        var = (C.c_int32 * max(n, 1))(*[int(v) for v in variables])
        out = (C.c_double * max(n, 1))()
        self._check(self.lib.lib.osb_probe(self.h, n, ijk, var, out))
        return np.array(list(out)[:n])

    def set_state_device(self, ptr: int):
        """State from device memory (interior layout, FP64), e.g. tensor.data_ptr()."""
        self._check(self.lib.lib.osb_set_state_device(self.h, C.c_void_p(ptr), LAYOUT_INTERIOR))

    def get_state_device(self, ptr: int):
        self._check(self.lib.lib.osb_get_state_device(self.h, C.c_void_p(ptr), LAYOUT_INTERIOR))

    def device_ptr(self, var: int) -> int:
        p = C.c_void_p()
        self._check(self.lib.lib.osb_state_device_ptr(self.h, int(var), C.byref(p)))
        return p.value

    def profile_enable(self, on: bool = True):
        """Start (and clear) or stop the library's per-kernel-family HIP-event timing."""
        self._check(self.lib.lib.osb_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        """{family: (total_ms, launches)} accumulated since profile_enable(True)."""
        out = {}
        i = 0
        while True:
            name = C.create_string_buffer(64)
            ms = C.c_double()
            n = C.c_int64()
            rc = self.lib.lib.osb_profile_read(self.h, i, name, 64, C.byref(ms), C.byref(n))
            if rc != OK:
                break
            out[name.value.decode()] = (ms.value, n.value)
            i += 1
        return out

    def describe(self) -> dict:
        """The context's kernel path and launch shapes (osb_describe)."""
        import json
        buf = C.create_string_buffer(4096)
        self._check(self.lib.lib.osb_describe(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_stream(self, stream_handle: int):
        self._check(self.lib.lib.osb_set_stream(self.h, C.c_void_p(stream_handle)))


class Batch:
    """An osb_batch: several instances advanced together (HIP library only)."""

    def __init__(self, lib: Library, descs: Sequence[ProblemDesc]):
        if not lib.has_batch:
            raise OSBError(E_UNSUPPORTED, f"{lib.path} has no batch entry points")
        self.lib = lib
        self.descs = list(descs)
        arr = (ProblemDescC * len(self.descs))(*[d.to_c() for d in self.descs])
        self._carr = arr
        h = C.c_void_p()
        rc = lib.lib.osb_batch_create(arr, len(self.descs), C.byref(h))
        if rc != OK:
            msg = lib.lib.osb_last_error(None)
            raise OSBError(rc, msg.decode() if msg else "")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.lib.osb_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != OK:
            msg = self.lib.lib.osb_batch_last_error(self.h)
            raise OSBError(rc, msg.decode() if msg else "")

    def set_state(self, i: int, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(self.lib.lib.osb_batch_set_state(self.h, int(i), q.reshape(-1), LAYOUT_INTERIOR))

    def get_state(self, i: int) -> np.ndarray:
        out = np.empty(self.descs[i].shape, dtype=np.float64)
        self._check(self.lib.lib.osb_batch_get_state(self.h, int(i), out.reshape(-1), LAYOUT_INTERIOR))
        return out

    def stats(self, i: int) -> StatsC:
        st = StatsC()
        st.struct_size = C.sizeof(StatsC)
        self._check(self.lib.lib.osb_batch_get_stats(self.h, int(i), C.byref(st)))
        return st

    def autotune(self):
        self._check(self.lib.lib.osb_batch_autotune(self.h))

    def describe(self) -> dict:
        import json
        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.lib.osb_batch_describe(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def step(self, nsteps: int = 1):
        self._check(self.lib.lib.osb_batch_step(self.h, int(nsteps)))

    def sync(self):
        self._check(self.lib.lib.osb_batch_sync(self.h))


class Pool:
    """An osb_pool: continuous batching of independent solves (HIP library only)."""

    def __init__(self, lib: Library, hbm_budget_bytes: int, max_solves: int = 64, nstreams: int = 4):
        if not lib.has_pool:
            raise OSBError(E_UNSUPPORTED, f"{lib.path} has no pool entry points")
        self.lib = lib
        h = C.c_void_p()
        rc = lib.lib.osb_pool_create(int(hbm_budget_bytes), int(max_solves), int(nstreams), C.byref(h))
        if rc != OK:
            msg = lib.lib.osb_last_error(None)
            raise OSBError(rc, msg.decode() if msg else "")
        self.h = h
        self.max_solves = max_solves
        self.descs = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.lib.osb_pool_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != OK:
            msg = self.lib.lib.osb_pool_last_error(self.h)
            raise OSBError(rc, msg.decode() if msg else "")

    def admit(self, desc: ProblemDesc, q0: np.ndarray, nsteps: int) -> Optional[int]:
        """Slot of the admitted solve, or None when slots or memory are short."""
        cd = desc.to_c()
        q0 = np.ascontiguousarray(q0, dtype=np.float64).reshape(-1)
        slot = C.c_int(-1)
        rc = self.lib.lib.osb_pool_admit(self.h, C.byref(cd), q0, int(nsteps), C.byref(slot))
        if rc == E_NOMEM:
            return None
        self._check(rc)
        self.descs[slot.value] = desc
        return slot.value

    def advance(self, chunk: int):
        self._check(self.lib.lib.osb_pool_advance(self.h, int(chunk)))

    def poll(self):
        buf = (C.c_int * self.max_solves)()
        n = C.c_int()
        self._check(self.lib.lib.osb_pool_poll(self.h, buf, self.max_solves, C.byref(n)))
        return [buf[i] for i in range(min(n.value, self.max_solves))]

    def fetch(self, slot: int):
        desc = self.descs[slot]
        out = np.empty(desc.shape, dtype=np.float64)
        st = StatsC()
        st.struct_size = C.sizeof(StatsC)
        self._check(self.lib.lib.osb_pool_fetch(self.h, int(slot), out.reshape(-1), C.byref(st)))
        return out, st

    def release(self, slot: int):
        self._check(self.lib.lib.osb_pool_release(self.h, int(slot)))
        self.descs.pop(slot, None)

    def usage(self):
        used = C.c_int64()
        running = C.c_int()
        done = C.c_int()
        self._check(self.lib.lib.osb_pool_usage(self.h, C.byref(used), C.byref(running), C.byref(done)))
        return used.value, running.value, done.value
