"""HIP back end for an OpenSBLI algorithm.

``HIPTarget(alg, ...)`` is called where a problem script called
``OPSC(alg)`` (opensbli/code_generation/opsc.py:408) followed by
``substitute_simulation_parameters(names, values)``
(opensbli/utilities/helperfunctions.py:150).  Instead of printing an OPS C
program it reads the block's schemes, boundary conditions and physics, checks
that the HIP kernels implement them, fills an ``osb_problem_desc`` and drives
``libosb_hip.so`` through the C ABI.

The block may be built from ``opensbli_amd.surface`` or from the reference's
own classes: only class names and the attributes named below are read.

Parameters come as the same name -> value mapping the reference pastes into
the generated C (``'Delta0block0': '2*M_PI/block0np0'``); values may be
numbers or C expressions over earlier names, evaluated in order with C
arithmetic (int/int divides as integers).
"""
from __future__ import annotations

import ast
import math
import operator
import os
import re
from typing import Callable, Dict, Mapping, Optional, Union

import numpy as np

from ... import native as nat


# --------------------------------------------------------------------------
# parameters: name -> number or C expression
# --------------------------------------------------------------------------
_FUNCS = {"ceil": math.ceil, "floor": math.floor, "sqrt": math.sqrt, "pow": math.pow, "exp": math.exp,
          "log": math.log, "sin": math.sin, "cos": math.cos, "tan": math.tan, "fabs": math.fabs,
          "abs": abs, "fmin": min, "fmax": max}
_CONSTS = {"M_PI": math.pi, "pi": math.pi}


def _c_div(a, b):
    if isinstance(a, int) and isinstance(b, int):
        q = abs(a) // abs(b)
        return q if (a >= 0) == (b >= 0) else -q
    return a / b


_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: _c_div,
           ast.Pow: operator.pow, ast.Mod: operator.mod}


def _eval_node(node, env):
    if isinstance(node, ast.Expression):
        return _eval_node(node.body, env)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return node.value
    if isinstance(node, ast.Name):
        if node.id in env:
            return env[node.id]
        if node.id in _CONSTS:
            return _CONSTS[node.id]
        raise KeyError(f"unknown name {node.id!r} in a simulation parameter")
    if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
        return _BINOPS[type(node.op)](_eval_node(node.left, env), _eval_node(node.right, env))
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval_node(node.operand, env)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS:
        return _FUNCS[node.func.id](*[_eval_node(a, env) for a in node.args])
    raise ValueError(f"unsupported expression in a simulation parameter: {ast.dump(node)}")


def evaluate_parameters(parameters: Mapping[str, Union[str, float, int]]) -> Dict[str, float]:
    """Evaluate ``substitute_simulation_parameters``-style values in order."""
    env: Dict[str, float] = {}
    for name, value in parameters.items():
        if isinstance(value, (int, float)):
            env[name] = value
        else:
            env[name] = _eval_node(ast.parse(str(value).strip(), mode="eval"), env)
    return env


# --------------------------------------------------------------------------
# plan: block -> osb_problem_desc
# --------------------------------------------------------------------------
def _kinds(obj) -> set:
    """Class names along the MRO (so subclasses such as StoreSome count as Central)."""
    return {k.__name__ for k in type(obj).__mro__}


_BC_IDS = {"PeriodicBC": nat.BC_PERIODIC, "ZeroGradientOutletBC": nat.BC_ZERO_GRADIENT,
           "InviscidWallBC": nat.BC_INVISCID_WALL, "SymmetryBC": nat.BC_SYMMETRY,
           "ExtrapolationBC": None, "DirichletBC": nat.BC_DIRICHLET}


def _conservative_names(ndim: int):
    return ["rho"] + [f"rhou{d}" for d in range(ndim)] + ["rhoE"]


def _equation_parts(eq):
    """(assigned name, right-hand side text) of one boundary equation: a SymPy
    / OpenSBLIEq equation (``lhs``/``rhs``) or the text ``Eq(DataObject(rho), d)``."""
    if hasattr(eq, "lhs") and hasattr(eq, "rhs"):
        # block datasets print as name_B0[x0, ...]; the variable is the name
        plain = lambda e: re.sub(r"\b([A-Za-z_]\w*?)_B\d+\[[^\]]*\]", r"\1", str(e))
        return plain(eq.lhs), plain(eq.rhs)
    node = ast.parse(str(eq).strip(), mode="eval").body
    if not (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == "Eq"
            and len(node.args) == 2):
        raise NotImplementedError(f"Dirichlet equation {eq!r}: expected Eq(lhs, rhs)")
    lhs = node.args[0]
    if (isinstance(lhs, ast.Call) and isinstance(lhs.func, ast.Name)
            and lhs.func.id in ("DataObject", "GridVariable") and len(lhs.args) == 1
            and isinstance(lhs.args[0], ast.Name)):
        name = lhs.args[0].id
    elif isinstance(lhs, ast.Name):
        name = lhs.id
    else:
        raise NotImplementedError(f"Dirichlet equation {eq!r}: unsupported left-hand side")
    return name, ast.unparse(node.args[1])


def dirichlet_values(bc, ndim: int, parameters: Mapping[str, float]):
    """Conservative values that DirichletBC ``bc`` (dirichlet.py:28-41) holds on
    its face, evaluated in order on the host with the simulation parameters in
    scope.  Every conservative variable must be assigned a constant."""
    names = _conservative_names(ndim)
    eqs = bc.equations
    if isinstance(eqs, Mapping):
        vals = {str(k): float(v) for k, v in eqs.items()}
    else:
        env = dict(parameters)
        vals = {}
        for eq in eqs:
            name, rhs = _equation_parts(eq)
            text = rhs.replace("DataObject(", "(").replace("GridVariable(", "(")
            try:
                env[name] = _eval_node(ast.parse(text, mode="eval"), env)
            except (KeyError, ValueError) as e:
                raise NotImplementedError(
                    f"DirichletBC(direction={bc.direction}, side={bc.side}): {name} = {rhs} is not a constant "
                    f"({e}); only constant boundary values are implemented") from None
            if name in names:
                vals[name] = float(env[name])
    missing = [m for m in names if m not in vals]
    if missing:
        raise NotImplementedError(
            f"DirichletBC(direction={bc.direction}, side={bc.side}) leaves {missing} unset; "
            "the HIP kernels hold every conservative variable")
    return [vals[m] for m in names]


def _averaging(scheme) -> int:
    avg = getattr(scheme, "averaging", None)
    if avg is None:
        return nat.AVG_SIMPLE            # shock_capturing.py:776-777
    kinds = _kinds(avg)
    if "RoeAverage" in kinds:
        return nat.AVG_ROE
    if "SimpleAverage" in kinds:
        return nat.AVG_SIMPLE
    raise NotImplementedError(f"averaging {type(avg).__name__} is not implemented by the HIP kernels")


def _viscous_model(physics) -> int:
    if physics is None:
        return nat.VISC_NONE
    visc = physics.get("viscosity") if isinstance(physics, Mapping) else getattr(physics, "viscosity", None)
    if visc in (None, "inviscid", "none"):
        return nat.VISC_NONE
    if visc == "constant":
        return nat.VISC_CONSTANT
    if visc in ("sutherland", "dynamic", "variable"):
        return nat.VISC_SUTHERLAND
    raise NotImplementedError(f"viscosity {visc!r} is not implemented by the HIP kernels")


def _convective_form(physics) -> str:
    if physics is None:
        return "Feiereisen"
    if isinstance(physics, Mapping):
        return physics.get("convective", physics.get("split_type", "Feiereisen"))
    return getattr(physics, "split_type", None) or type(getattr(physics, "split", None)).__name__


def plan_problem(alg_or_block, parameters: Mapping, physics=None, precision: str = "fp64",
                 cfl: Optional[float] = None) -> nat.ProblemDesc:
    """Select the kernel plan for one block and return its descriptor.

    ``physics``: an ``NS_Split`` object (its ``split_type`` and ``viscosity``)
    or a dict ``{"convective": "Feiereisen"|"Skew", "viscosity":
    "constant"|"sutherland"|None}``; it supplies what the reference keeps in
    the SymPy equations.  Combinations the kernels do not implement raise
    ``NotImplementedError`` naming the offending object."""
    block = alg_or_block
    if hasattr(alg_or_block, "blocks"):
        if len(alg_or_block.blocks) != 1:
            raise NotImplementedError("the HIP back end drives one block per algorithm")
        block = alg_or_block.blocks[0]
    p = evaluate_parameters(parameters)
    ndim = int(block.ndim)
    n = [int(p[f"block0np{d}"]) for d in range(ndim)]
    delta = [float(p[f"Delta{d}block0"]) for d in range(ndim)]

    shock, central, time = None, None, None
    for sch in block.discretisation_schemes.values():
        kinds = _kinds(sch)
        if kinds & {"LFWeno", "LFTeno", "HLLCWeno"}:
            if shock is not None:
                raise NotImplementedError("one shock-capturing scheme per block")
            shock = sch
        elif "Central" in kinds:
            central = sch
        elif kinds & {"RungeKuttaLS", "RungeKutta"}:
            time = sch
        else:
            raise NotImplementedError(f"scheme {type(sch).__name__} is not implemented by the HIP kernels")
    if time is None:
        raise NotImplementedError("no temporal scheme on the block")

    desc = dict(ndim=ndim, n=n, delta=delta)
    if shock is not None:
        kinds = _kinds(shock)
        ft = getattr(shock, "flux_type", None) or ("HLLC" if "HLLCWeno" in kinds else "LLF")   # shock_capturing.py:780-781
        if "HLLCWeno" in kinds:
            if ft not in ("HLLC", "HLLC-LM"):
                raise NotImplementedError(f"HLLCWeno(flux_type={ft!r}): HLLC and HLLC-LM are implemented")
            desc["flux_type"] = nat.FLUX_HLLC_LM if ft == "HLLC-LM" else nat.FLUX_HLLC
        elif ft not in ("LLF", "GLF"):
            raise NotImplementedError(f"{type(shock).__name__}(flux_type={ft!r}): LLF and GLF are implemented")
        else:
            desc["flux_type"] = nat.FLUX_GLF if ft == "GLF" else nat.FLUX_LLF
        if kinds & {"LFWeno", "HLLCWeno"}:
            if shock.order != 5 or str(getattr(shock, "formulation", "JS")).upper() != "Z":
                raise NotImplementedError(
                    f"LFWeno(order={shock.order}, formulation={shock.formulation!r}): only order 5, 'Z'")
            desc.update(conv_scheme=nat.CONV_WENO_Z, order=5)
        else:
            if shock.order not in (5, 6) or getattr(shock, "formulation", None) is not None:
                raise NotImplementedError(f"LFTeno(order={shock.order}): orders 5 and 6 without a sensor")
            desc.update(conv_scheme=nat.CONV_TENO, order=int(shock.order))
            if "TENO_CT" in p:
                desc["teno_ct"] = float(p["TENO_CT"])
            if "eps" in p:
                desc["teno_eps"] = float(p["eps"])
        desc["averaging"] = _averaging(shock)
        if "HLLCWeno" in kinds and getattr(shock, "averaging", None) is None:
            desc["averaging"] = nat.AVG_ROE        # HLLCCharacteristic's default (shock_capturing.py:1040-1041)
        desc["viscous"] = _viscous_model(physics)
        if desc["viscous"] != nat.VISC_NONE and (central is None or central.order != 4):
            raise NotImplementedError("viscous terms need a Central(4) scheme on the block")
    else:
        if central is None:
            raise NotImplementedError("no spatial scheme on the block")
        if central.order != 4:
            raise NotImplementedError(f"{type(central).__name__}(order={central.order}): only order 4")
        form = _convective_form(physics)
        if form == "Feiereisen":
            desc["conv_scheme"] = nat.CONV_CENTRAL_FEIEREISEN
        elif form in ("Skew", "skew"):
            desc["conv_scheme"] = nat.CONV_CENTRAL_SKEW
        elif form == "KGP":
            if isinstance(physics, Mapping):
                ef = physics.get("energy_formulation", "none")
            else:        # the mirror keeps it on NS_Split, the reference on its split object
                ef = getattr(physics, "energy_formulation", None) or \
                    getattr(getattr(physics, "split", None), "energy_formulation", "none")
            if ef != "enthalpy":
                raise NotImplementedError(f"NS_Split('KGP', energy_formulation={ef!r}): the enthalpy formulation is implemented")
            desc["conv_scheme"] = nat.CONV_CENTRAL_KGP
        else:
            raise NotImplementedError(f"convective split form {form!r}: Feiereisen, KGP (enthalpy) and Skew are implemented")
        desc.update(order=4, viscous=_viscous_model(physics))

    tk = _kinds(time)
    if "RungeKuttaLS" in tk:
        desc["time_scheme"] = nat.TIME_RKLS3
        a = list(getattr(time, "rk_a", []))
        b = list(getattr(time, "rk_b", []))
        williamson = time.order == 3 and getattr(time, "formulation", None) != "SSP"
        if not williamson:
            if not a:
                raise NotImplementedError(f"RungeKuttaLS(order={time.order}) without rk_a/rk_b coefficients")
            desc.update(rk_a=a, rk_b=b)
    else:
        if time.order != 3:
            raise NotImplementedError(f"RungeKutta(order={time.order}): only order 3")
        desc["time_scheme"] = nat.TIME_RK3

    bc = [[0, 0] for _ in range(ndim)]
    bc_value = {}
    for pair in block.boundary_types:
        for b in pair:
            name = next((k for k in _kinds(b) if k in _BC_IDS), None)
            if name is None:
                raise NotImplementedError(f"boundary condition {type(b).__name__} is not implemented by the HIP kernels")
            if name == "ExtrapolationBC":
                if b.order not in (0, 1):
                    raise NotImplementedError(f"ExtrapolationBC(order={b.order}): orders 0 and 1")
                bc[b.direction][b.side] = nat.BC_EXTRAPOLATION0 if b.order == 0 else nat.BC_EXTRAPOLATION1
            else:
                bc[b.direction][b.side] = _BC_IDS[name]
            if name == "DirichletBC":
                bc_value[(b.direction, b.side)] = dirichlet_values(b, ndim, p)
            if name in ("ExtrapolationBC", "DirichletBC") and getattr(b, "modification_scheme", None) is not None \
                    and type(b.modification_scheme).__name__ != "Carpenter":
                raise NotImplementedError(f"{name}(scheme={type(b.modification_scheme).__name__}): the kernels "
                                          "implement the default Carpenter closure")
    if any(v == 0 for pair in bc for v in pair):
        raise ValueError("every direction needs a boundary condition on both sides")

    # non-linear filters among the equation classes (WENO_filter.py:406-443)
    filters = [ec for ec in getattr(block, "list_of_equation_classes", []) if type(ec).__name__.endswith("Filter")]
    for f in filters:
        if type(f).__name__ != "WENOFilter":
            raise NotImplementedError(f"{type(f).__name__} is not implemented by the HIP kernels")
        if f.order not in (3, 5) or str(f.formulation).upper() != "Z" or f.flux_type != "LLF" \
                or f.metrics is not None or f.species is not None:
            raise NotImplementedError(
                f"WENOFilter(order={f.order}, formulation={f.formulation!r}, flux_type={f.flux_type!r}, "
                f"metrics={f.metrics!r}): orders 3 and 5, 'Z', 'LLF', uniform mesh only")
        if f.optimize and ndim < 2:
            raise NotImplementedError("WENOFilter(optimize=True): the Ducros sensor needs ndim >= 2")
    if len(filters) > 1:
        raise NotImplementedError("one WENO filter per block")
    if filters:
        if desc.get("flux_type") == nat.FLUX_GLF:
            raise NotImplementedError("a WENO filter over a GLF base scheme: the kernels pair the filter with LLF")
        desc["shock_filter"] = nat.FILTER_WENOZ
        f = filters[0]
        desc["filter_order"] = int(f.order)
        if f.optimize:                       # Ducros hybrid selection (WENO_filter.py:455-458, 561-584)
            desc["filter_sensor"] = nat.SENSOR_DUCROS
            desc["ducros_check"] = float(p.get("Ducros_check", 0.05))
            desc["ducros_select"] = float(p.get("Ducros_select", 0.05))
        for d in range(ndim):
            key = f"inv_rfact{d}_block0"         # the filter's 1/Delta (opensblifunctions.py:577-591)
            if key in p and not math.isclose(float(p[key]), 1.0 / delta[d], rel_tol=1e-12):
                raise NotImplementedError(f"{key} = {p[key]}: the filter kernels scale by 1/Delta{d}block0")
    desc["bc"] = bc
    desc["bc_value"] = bc_value

    for key, name in (("gama", "gama"), ("minf", "Minf"), ("re", "Re"), ("pr", "Pr"),
                      ("shock_filter_control", "shock_filter_control")):   # shock_capturing.py:1014-1016
        if name in p:
            desc[key] = float(p[name])
    if cfl is not None:
        desc["cfl"] = float(cfl)
    elif "cfl" in p:
        desc["cfl"] = float(p["cfl"])
    if "dt" in p:
        desc["dt"] = float(p["dt"])
    desc["precision"] = nat.PREC_MIXED if precision == "mixed" else nat.PREC_FP64
    return nat.ProblemDesc(**desc)


def _io_kwargs(io) -> Dict:
    """iohdf5 keyword choices: the mirror's ``kwargs`` (lower-cased like
    io_hdf5.py:33-38) or the reference object's."""
    kw = getattr(io, "kwargs", None) or {}
    return {str(k).lower(): (v.lower() if isinstance(v, str) else v) for k, v in kw.items()}


def _io_filename(io) -> str:
    """Output file of a writer outside the loop (io_hdf5.py:190-206)."""
    name = str(_io_kwargs(io).get("name") or "opensbli_output")
    return name if name.endswith(".h5") else f"{name}.h5"


def _save_every(io) -> int:
    v = getattr(io, "save_every", None)
    v = getattr(v, "_value", v)           # the reference keeps a ConstantObject
    return int(v) if v else 0


def initial_fields(block, parameters: Mapping[str, float], desc: nat.ProblemDesc) -> np.ndarray:
    """Evaluate the block's GridBasedInitialisation equations in order over the
    grid points [0, N) on the host (opensbli/initialisation/gridbasedinit.py:16-86)
    and return the conservative state (nv, n2, n1, n0).  Grid indexes enter as
    idx0..idx2 (the reference's idx[d]); simulation parameters such as gama and
    Delta0block0 are in scope.  The halo points the reference also fills are
    rewritten by the boundary conditions before the first stage, so only the
    interior is evaluated."""
    import sympy as sp
    eqs = []
    for ec in getattr(block, "list_of_equation_classes", []):
        if "GridBasedInitialisation" in _kinds(ec):
            eqs += list(ec.equations)
    if not eqs:
        raise ValueError("the block has no GridBasedInitialisation equations")
    from ... import surface
    local = {"block": block, "DataObject": surface.DataObject, "GridVariable": surface.GridVariable, "Eq": sp.Eq}
    n = list(desc.n) + [1] * (3 - len(desc.n))
    shape = (n[2], n[1], n[0])
    grids = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    env: Dict[str, object] = {k: v for k, v in parameters.items() if isinstance(v, (int, float))}
    for d in range(desc.ndim):
        env[f"idx{d}"] = grids[2 - d].astype(np.float64)
    plain = lambda e: re.sub(r"\b([A-Za-z_]\w*?)_B\d+\[[^\]]*\]", r"\1", str(e))
    for eq in eqs:
        if isinstance(eq, str):
            eq = sp.parse_expr(eq, local_dict=local)
        rhs = eq.rhs
        # the reference's grid index idx[d] (an Indexed) -> idx<d>
        rhs = rhs.xreplace({a: sp.Symbol(f"idx{int(a.indices[0])}") for a in rhs.atoms(sp.Indexed)
                            if str(a.base.label).startswith("idx")})
        syms = sorted(rhs.free_symbols, key=str)
        names = [plain(s_) for s_ in syms]
        missing = [m for m in names if m not in env]
        if missing:
            raise NotImplementedError(f"initial condition {eq}: unknown names {missing}")
        f = sp.lambdify(syms, rhs, modules="numpy")
        env[plain(eq.lhs)] = f(*[env[m] for m in names])
    q = np.empty(desc.shape, dtype=np.float64)
    for v, nm in enumerate(_conservative_names(desc.ndim)):
        if nm not in env:
            raise NotImplementedError(f"the initial equations do not set {nm} (conservative variables are required)")
        q[v] = np.broadcast_to(np.asarray(env[nm], dtype=np.float64), shape)
    return q


# --------------------------------------------------------------------------
# the target
# --------------------------------------------------------------------------
class HIPTarget:
    """Runs an OpenSBLI algorithm on the HIP kernels.

    >>> target = HIPTarget(alg, parameters=simulation_parameters, physics=NS)
    >>> target.initialise(q0)          # array (nv, n2, n1, n0), or f(x0, x1, x2)
    >>> target.run()                   # niter steps, state stays on the GPU
    >>> q = target.state()

    ``library`` defaults to ``libosb_hip.so``; any library exporting the same
    C ABI can be passed (the tests pass the CPU oracle)."""

    def __init__(self, alg, precision: str = "fp64", parameters: Optional[Mapping] = None, physics=None,
                 cfl: Optional[float] = None, library: Optional[nat.Library] = None, output_dir: str = "."):
        self.parameters = evaluate_parameters(parameters or {})
        self.desc = plan_problem(alg, parameters or {}, physics=physics, precision=precision, cfl=cfl)
        self.niter = int(self.parameters.get("niter", 0))
        self.lib = library if library is not None else nat.load_hip()
        self.ctx = self.lib.context(self.desc)
        self.output_dir = output_dir
        block = alg.blocks[0] if hasattr(alg, "blocks") else alg
        self.block = block
        self.block_number = int(getattr(block, "blocknumber", 0) or 0)
        ios = list(getattr(block, "InputOutput", []) or [])
        self.writers = [io for io in ios if str(_io_kwargs(io).get("iotype", "write")).lower() == "write"]
        self.restart_file = None
        for io in ios:
            kw = _io_kwargs(io)
            if str(kw.get("iotype", "")).lower() == "read" and str(kw.get("restart_simulation", "")).lower() in (
                    "true", "1", "yes"):
                self.restart_file = os.path.join(output_dir, "restart.h5")      # opsc.py:962-977
        self.start_iter = 0
        self.monitor = getattr(alg, "simulation_monitor", None)
        self._probes = None
        if self.monitor is not None:
            names = _conservative_names(self.desc.ndim)
            pts, var = [], []
            for a, loc in zip(self.monitor.arrays, self.monitor.probe_locations):
                if not isinstance(loc, tuple):
                    raise NotImplementedError(f"SimulationMonitor scalar monitor {a!r}: only point probes are implemented")
                nm = re.sub(r"_B\d+$", "", str(a))
                if nm not in names:
                    raise NotImplementedError(f"SimulationMonitor array {a!r}: conservative variables only ({names})")
                pts.append(tuple(int(x) for x in loc))
                var.append(names.index(nm))
            self._probes = (pts, var)

    # grid coordinates of the interior points, x_d = Delta_d * i_d
    def coordinates(self):
        n = list(self.desc.n) + [1] * (3 - len(self.desc.n))
        dl = list(self.desc.delta) + [1.0] * (3 - len(self.desc.delta))
        axes = [dl[d] * np.arange(n[d], dtype=np.float64) for d in range(3)]
        z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        return x, y, z

    def initialise(self, q0: Union[np.ndarray, Callable, None] = None):
        """Initial state: an array (nv, n2, n1, n0), a function of the grid
        coordinates, or (None) the block's GridBasedInitialisation equations."""
        if q0 is None:
            q0 = initial_fields(self.block, self.parameters, self.desc)
        elif callable(q0):
            q0 = np.asarray(q0(*self.coordinates()), dtype=np.float64)
        self.ctx.set_state(np.asarray(q0, dtype=np.float64).reshape(self.desc.shape))
        return self

    def restart(self, path: Optional[str] = None):
        """State, start iteration and simulation time from an HDF5 restart file."""
        from ... import io_hdf5
        self.start_iter, _ = io_hdf5.restart(self.ctx, path or self.restart_file or "restart.h5", self.block_number)
        return self

    def write_hdf5(self, path: str, iteration: Optional[int] = None):
        """The block's conservative variables and the simulation constants in
        the OPS HDF5 layout (ops_fetch_dat_hdf5_file + write_constants)."""
        from ... import io_hdf5
        consts = {k: v for k, v in self.parameters.items() if isinstance(v, (int, float)) and k not in ("iter",)}
        ints = {k for k, v in consts.items() if isinstance(v, int) and not isinstance(v, bool)}
        io_hdf5.write_state(path, self.ctx, consts, ints, iteration=iteration, block_number=self.block_number)

    def run(self, niter: Optional[int] = None):
        """niter steps; the block's iohdf5 writers run where the generated time
        loop places them (algorithm.py:476-485): with save_every after loop
        index it when (it + 1) % save_every == 0 or it == 0, as
        opensbli_output_%06d.h5 % (it + 1), and at the end as
        opensbli_output.h5 (or the writer's name)."""
        steps = self.niter if niter is None else int(niter)
        if self.restart_file is not None and self.ctx.stats().iter == 0:
            self.restart()
        for w in self.writers:                       # position='init': before the time loop only
            if _io_kwargs(w).get("position") == "init" and self.ctx.stats().iter == 0:
                self.write_hdf5(os.path.join(self.output_dir, _io_filename(w)))
        every = sorted({_save_every(w) for w in self.writers if _save_every(w)})
        mon = self.monitor
        periods = every + ([mon.frequency] if mon is not None else [])
        it = int(self.ctx.stats().iter)              # loop index of the next iteration
        stop = it + steps
        if not periods:
            self.ctx.step(steps)
        while periods and it < stop:
            # next loop index j >= it with in-loop work: (j + 1) % period == 0, or j == 0
            j = min([stop - 1, 0 if it == 0 else stop] + [-(-(it + 1) // e) * e - 1 for e in periods])
            self.ctx.step(j - it + 1)
            if mon is not None and (j == 0 or (j + 1) % mon.frequency == 0):
                self._monitor(j)                     # the monitor precedes the in-loop IO (algorithm.py:452-485)
            if every and (j == 0 or any((j + 1) % e == 0 for e in every)):
                self.write_hdf5(os.path.join(self.output_dir, "opensbli_output_%06d.h5" % (j + 1)), iteration=j)
            it = j + 1
        for w in self.writers:
            if _io_kwargs(w).get("position") != "init":
                self.write_hdf5(os.path.join(self.output_dir, _io_filename(w)))
        return self

    def _monitor(self, j: int):
        """One monitor line after loop index j (simulation_monitors.py:205-257):
        iteration j + 1, the simulation time, then each probe, in %.<p>e; the
        header first when j == 0.  NaN_check first runs ops_NaNcheck."""
        mon = self.monitor
        if mon.NaN_check is not None:
            rc = self.ctx.nan_check()
            if rc:
                raise nat.OSBError(rc, f"NaN check failed after iteration {j + 1}")
        pts, var = self._probes
        vals = self.ctx.probe(pts, var) if pts else []
        p = mon.fp_precision
        lines = []
        if j == 0:
            heads = ["Iteration", "Time"] + [f"{re.sub(r'_B[0-9]+$', '', str(a))}_B{self.block_number}{loc}"
                                             for a, loc in zip(mon.arrays, mon.probe_locations)]
            lines.append(", ".join(heads))
        t = self.ctx.stats().time
        lines.append(", ".join([str(j + 1), f"{t:.{p}e}"] + [f"{v:.{p}e}" for v in vals]))
        if mon.output_file:
            path = os.path.join(self.output_dir, f"block{self.block_number}_{mon.output_file}")
            with open(path, "a") as f:
                f.write("\n".join(lines) + "\n")
        else:
            print("\n".join(lines), flush=True)

    def state(self) -> np.ndarray:
        return self.ctx.get_state()

    def stats(self):
        return self.ctx.stats()

    def close(self):
        self.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
