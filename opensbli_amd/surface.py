"""OpenSBLI problem-definition objects understood by the HIP back end.

The class names, constructor arguments and the attributes the back end reads
follow the reference, so a problem script written against OpenSBLI selects
the same kernels here.  Each class records the choices a script makes; it
does not build SymPy equations.  ``code_generation.hip.HIPTarget`` turns a
block built from these objects -- or from the reference's own objects, whose
class names and attributes it reads the same way -- into an
``osb_problem_desc``.

Only the configurations that the HIP kernels implement are accepted by the
back end; everything else raises ``NotImplementedError`` there, naming the
offending object.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence


# --------------------------------------------------------------------------
# characteristic averaging (opensbli/schemes/spatial/averaging.py:16-124)
# --------------------------------------------------------------------------
class Averaging:
    def __init__(self, locations: Sequence[int]):
        self.locations = list(locations)


class SimpleAverage(Averaging):
    """Arithmetic mean of the states at the two locations (averaging.py:31)."""


class RoeAverage(Averaging):
    """Roe average (averaging.py:65)."""

    def __init__(self, locations: Sequence[int], physics=None):
        super().__init__(locations)
        self.physics = physics


# --------------------------------------------------------------------------
# spatial schemes
# --------------------------------------------------------------------------
class Scheme:
    def __init__(self, name: str, order: int):
        self.name = name
        self.order = order


class LFWeno(Scheme):
    """Local Lax-Friedrichs characteristic WENO (opensbli/schemes/spatial/weno.py:475)."""

    def __init__(self, order, physics=None, averaging=None, shock_filter=None, formulation="JS",
                 conservative=True, flux_type="LLF", flux_split=True, species=None):
        if flux_type not in ("LLF", "GLF"):
            raise ValueError(f"flux_type must be 'LLF' or 'GLF' for Lax-Friedrichs splitting, not {flux_type!r}")
        if order % 2 == 0:
            raise ValueError(f"WENO schemes have odd orders; order {order} is not available")
        super().__init__("WenoDerivative", order)
        self.physics = physics
        self.averaging = averaging
        self.shock_filter = shock_filter
        self.formulation = formulation
        self.conservative = conservative
        self.flux_type = flux_type
        self.flux_split = flux_split
        self.species = species


class HLLCWeno(Scheme):
    """HLLC Riemann fluxes of WENO-reconstructed characteristic states
    (opensbli/schemes/spatial/weno.py:586, shock_capturing.py:1031);
    flux_type 'HLLC' or 'HLLC-LM' (low-Mach variant).  Without an averaging
    the characteristic decomposition uses RoeAverage (shock_capturing.py:1040-1041)."""

    def __init__(self, order, physics=None, averaging=None, shock_filter=None, formulation="JS",
                 conservative=True, flux_type="HLLC", flux_split=False):
        if order % 2 == 0:
            raise ValueError(f"WENO schemes have odd orders; order {order} is not available")
        if flux_type not in ("HLLC", "HLLC-LM"):
            raise ValueError(f"flux_type must be 'HLLC' or 'HLLC-LM' for the HLLC Riemann flux, not {flux_type!r}")
        super().__init__("WenoDerivative", order)
        self.physics = physics
        self.averaging = averaging
        self.shock_filter = shock_filter
        self.formulation = formulation
        self.conservative = conservative
        self.flux_type = flux_type
        self.flux_split = False


class LFTeno(Scheme):
    """Local Lax-Friedrichs characteristic TENO (opensbli/schemes/spatial/teno.py:480)."""

    def __init__(self, order, formulation=None, physics=None, averaging=None, sensor=None, store_sensor=False,
                 conservative=True, flux_type="LLF", flux_split=True, passive_scalar=False):
        if flux_type not in ("LLF", "GLF"):
            raise ValueError(f"flux_type must be 'LLF' or 'GLF' for Lax-Friedrichs splitting, not {flux_type!r}")
        super().__init__("TenoDerivative", order)
        self.formulation = formulation
        self.physics = physics
        self.averaging = averaging
        self.sensor_array = sensor
        self.store_sensor = store_sensor
        self.conservative = conservative
        self.flux_type = flux_type
        self.flux_split = flux_split
        self.passive_scalar = passive_scalar


class Central(Scheme):
    """Central differences (opensbli/schemes/spatial/scheme.py:64)."""

    def __init__(self, order):
        super().__init__("CentralDerivative", order)
        self.points = list(range(-(order // 2), order // 2 + 1))


class StoreSome(Central):
    """Central differences storing the derivatives of some functions
    (opensbli/optimisations/StoreSome.py:19)."""

    def __init__(self, order, der_fns_to_store, merged=True, group_stored=False, level=1):
        super().__init__(order)
        self.fns = der_fns_to_store
        self.merged = merged
        self.group_stored = group_stored


# --------------------------------------------------------------------------
# temporal schemes
# --------------------------------------------------------------------------
class RungeKuttaLS(Scheme):
    """2N-storage Runge-Kutta (opensbli/schemes/temporal/rk_LS.py:17).

    ``rk_a``/``rk_b`` are the stage coefficients of rk_LS.py:97-132: the
    Williamson 3-stage scheme, Ruuth's SSP 3rd-order schemes with 3, 4 or 5
    stages, and the Carpenter-Kennedy 4th-order 5-stage scheme.  The values
    are published coefficients and must be the same numbers."""

    def __init__(self, order, stages=None, formulation=None):
        super().__init__("RungeKutta", order)
        self.formulation = formulation
        if order == 3:
            if formulation == "SSP":
                n = 3 if stages is None else stages
                tables = {
                    3: ([0.0, -2.915493957701923, 0.0],
                        [0.924574112262461, 0.287712943868770, 0.626538293270800]),
                    4: ([0.0, -0.449336503268844, 0.0, -4.661555711601366],
                        [1.086620745813428, 0.854115548251602, -1.576604558206099, -0.278475500113052]),
                    5: ([0.0, -4.344339134485095, 0.0, -3.770024161386381, -0.046347284573284],
                        [0.713497331193829, 0.133505249805329, 0.713497331193829, 0.149579395628565,
                         0.384471116121269]),
                }
                if n not in tables:
                    raise ValueError(f"third-order SSP low-storage Runge-Kutta has 3, 4 or 5 stages, not {n}")
                self.rk_a, self.rk_b = tables[n]
            else:
                n = 3
                self.rk_a = [0.0, -5.0 / 9.0, -153.0 / 128.0]
                self.rk_b = [1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0]
        elif order == 4:
            n = 5
            self.rk_a = [0.0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257]
            self.rk_b = [0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681]
        else:
            raise NotImplementedError(f"low-storage Runge-Kutta of order {order}: orders 3 and 4 are available")
        self.stages = n
        self.n_stages = n

    @property
    def is_default(self) -> bool:
        """True for the Williamson 3-stage scheme (the descriptor's default)."""
        return self.order == 3 and self.formulation != "SSP"


class RungeKutta(Scheme):
    """Classic 3-stage RK3 with a saved q (opensbli/schemes/temporal/rk_sbli.py:16)."""

    def __init__(self, order, constant_dt=None):
        if order != 3:
            raise NotImplementedError(f"classic Runge-Kutta of order {order}: order 3 is available")
        super().__init__("RungeKutta", order)
        self.constant_dt = constant_dt


# --------------------------------------------------------------------------
# boundary conditions (opensbli/core/boundary_conditions/)
# --------------------------------------------------------------------------
class BoundaryConditionBase:
    def __init__(self, direction, side, plane=True):
        self.direction = direction
        self.side = side
        self.full_plane = bool(plane)


class PeriodicBC(BoundaryConditionBase):
    """periodic.py:7"""

    def __init__(self, direction, side, halos=None, full_depth=False, corners=True, plane=True):
        super().__init__(direction, side, plane)
        self.halos = halos
        self.corners = corners


class ZeroGradientOutletBC(BoundaryConditionBase):
    """zero_gradient_outlet.py:5"""

    def __init__(self, direction, side, scheme=None, plane=True):
        super().__init__(direction, side, plane)
        self.scheme = scheme


class InviscidWallBC(BoundaryConditionBase):
    """inviscid_wall.py:6"""

    def __init__(self, direction, side, scheme=None, plane=True, include_corner_halos=True):
        super().__init__(direction, side, plane)
        self.scheme = scheme


class SymmetryBC(BoundaryConditionBase):
    """symmetry.py:6"""

    def __init__(self, direction, side, scheme=None, plane=True):
        super().__init__(direction, side, plane)
        self.scheme = scheme


class Carpenter:
    """opensbli/core/boundary_conditions/Carpenter_scheme.py:6 -- the 4th-order
    one-sided closure the central derivatives switch to next to extrapolation
    and Dirichlet faces (the default ``modification_scheme`` of those BCs)."""


class ExtrapolationBC(BoundaryConditionBase):
    """extrapolation.py:8.  order 0: the boundary point and the halos take the
    first interior point; order 1: linear extrapolation into the halos."""

    def __init__(self, direction, side, order, scheme=None, plane=True):
        super().__init__(direction, side, plane)
        if order not in (0, 1):
            raise ValueError(f"extrapolation order {order}: orders 0 (constant) and 1 (linear) are available")
        self.order = order
        self.modification_scheme = scheme if scheme else Carpenter()


class DirichletBC(BoundaryConditionBase):
    """dirichlet.py:6.  ``equations`` as the reference's problem scripts write
    them -- ``"Eq(GridVariable(d), 1.0)"`` ... ``"Eq(DataObject(rho), d)"``, as
    text or as SymPy equations -- or a mapping {conservative name: value}.  The
    HIP back end evaluates them once on the host (simulation parameters such as
    ``gama`` in scope) and holds every conservative variable at the value on the
    boundary point and this side's halos; values that depend on the grid
    position are refused there."""

    def __init__(self, direction, side, equations, scheme=None, plane=True):
        super().__init__(direction, side, plane)
        self.equations = equations
        self.modification_scheme = scheme if scheme else Carpenter()


# --------------------------------------------------------------------------
# equations, physics and the block
# --------------------------------------------------------------------------
class NS_Split:
    """Split forms of the Navier-Stokes convective terms
    (opensbli/physical_models/split_forms.py:295).  ``viscosity`` is 'constant'
    (tau = (1/Re)(...)), 'inviscid' or None (no diffusive terms), or a
    variable-viscosity choice ('dynamic', 'sutherland'): tau = (mu/Re)(...)
    with the Sutherland law mu = 1.4042 T^1.5/(T + 0.40417)."""

    def __init__(self, split_type, ndim, constants=(), coordinate_symbol="x", conservative=True, viscosity=None,
                 energy_formulation="none", debug=False):
        known = ("Divergence", "Blaisdell", "Feiereisen", "Jameson", "Kok", "KGP", "KEEP")
        if split_type not in known:
            raise NotImplementedError(f"split form {split_type!r} is not one of {', '.join(known)}")
        self.split_type = split_type
        self.ndim = ndim
        self.viscosity = viscosity
        self.inviscid = viscosity == "inviscid"
        self.conservative = conservative
        self.energy_formulation = energy_formulation
        self.mass = ("NS_Split", split_type, "mass")
        self.momentum = ("NS_Split", split_type, "momentum")
        self.energy = ("NS_Split", split_type, "energy")


class EinsteinEquation:
    """Stand-in for the index expander: returns the equation text unchanged."""

    def expand(self, equation, ndim, coordinate_symbol, substitutions, constants):
        return equation


class _EquationClass:
    def __init__(self):
        self.equations: List = []

    def add_equations(self, eqs):
        self.equations += list(eqs) if isinstance(eqs, (list, tuple)) else [eqs]


class SimulationEquations(_EquationClass):
    """opensbli/equation_types/opensbliequations.py"""


class ConstituentRelations(_EquationClass):
    """opensbli/equation_types/opensbliequations.py:491"""


class GridBasedInitialisation(_EquationClass):
    """opensbli/initialisation/gridbasedinit.py:16.  The equations as a
    problem script writes them (SymPy equations from ``parse_expr`` with
    ``block``, ``DataObject`` and ``GridVariable`` in the local dict, or the
    equation strings themselves); the HIP back end evaluates them in order
    over the grid on the host (``HIPTarget.initialise()`` without an array)."""


def _symbol(name):
    try:
        import sympy
    except ImportError as e:                          # pragma: no cover - SymPy is in the image
        raise ImportError("equation objects need SymPy") from e
    return sympy.Symbol(str(name))


def DataObject(name):
    """A named grid array in equations (opensbli/core/opensbliobjects.py);
    a SymPy symbol of that name here."""
    return _symbol(name)


def GridVariable(name):
    """A point-local temporary in equations (opensbli/core/grid.py:214); a
    SymPy symbol of that name here."""
    return _symbol(name)


class iohdf5:
    """opensbli/core/io_hdf5.py:22 -- records the arrays a script asks to write."""

    def __init__(self, arrays=None, save_every=None, **kwargs):
        self.arrays = list(arrays or [])
        self.save_every = save_every
        self.kwargs = {k.lower(): (v.lower() if isinstance(v, str) else v) for k, v in kwargs.items()} \
            or {"iotype": "write"}

    def add_arrays(self, arrays):
        self.arrays += list(arrays)


class Double:
    """SimulationDataType choice: FP64 everywhere."""


class FloatC:
    """SimulationDataType choice of the reference for FP32 storage; the HIP
    back end offers instead the FP32-flux/FP64-state 'mixed' precision."""


class SimulationDataType:
    """opensbli/core/datatypes.py:3"""
    dtype_ = Double

    @staticmethod
    def set_datatype(types):
        SimulationDataType.dtype_ = types

    @staticmethod
    def dtype():
        return SimulationDataType.dtype_


class SimulationBlock:
    """One structured block (opensbli/core/block.py:83)."""

    def __init__(self, ndim, block_number=0):
        self.ndim = ndim
        self.blocknumber = block_number
        self.block_number = block_number
        self.boundary_types: List = []
        self.discretisation_schemes: Dict[str, Scheme] = {}
        self.list_of_equation_classes: List = []
        self.InputOutput: List = []

    @property
    def deltas(self):
        """Grid spacings Delta{d}block{b} (opensbli/core/grid.py:155)."""
        return [_symbol(f"Delta{d}block{self.blocknumber}") for d in range(self.ndim)]

    @property
    def grid_indexes(self):
        """Grid index of each direction (opensbli/core/grid.py:176), idx0..idx2."""
        return [_symbol(f"idx{d}") for d in range(self.ndim)]

    def set_block_boundaries(self, bclist):
        flat = _flatten(bclist)
        counts = [0, 0, 0]
        for bc in flat:
            counts[bc.direction] += 1
        for d in range(self.ndim):
            if counts[d] != 2:
                raise AssertionError(f"direction {d} needs two boundary conditions, got {counts[d]}")
        it = iter(flat)
        self.boundary_types = list(zip(it, it))

    def set_discretisation_schemes(self, schemes: Dict[str, Scheme]):
        self.discretisation_schemes = schemes

    def set_equations(self, list_of_equations):
        self.list_of_equation_classes += list(list_of_equations)

    def setio(self, list_of_io):
        self.InputOutput += _flatten([list_of_io])

    def discretise(self):
        return None


class WENOFilter:
    """opensbli/filters/WENO_filter.py:406 -- the non-linear WENO filter applied
    after each full time step of a non-dissipative base scheme: the
    dissipative part of a WENO-Z reconstruction in characteristic space (its
    interface flux less the 6th-order central interpolation of the
    characteristic fluxes) scaled by the WENO weight sensor, with Roe
    averaging, and q <- q + dt * (filter residual).  ``equation_classes`` goes
    to ``block.set_equations`` as in the reference.  The HIP kernels
    implement order 5, formulation 'Z', flux_type 'LLF', uniform meshes
    (metrics=None) and optimize=False (no Ducros hybrid selection)."""

    def __init__(self, block, order, metrics=None, flux_type="LLF", airfoil=False, formulation="Z", optimize=False,
                 species=None):
        self.block = block
        self.order = int(order)
        self.metrics = metrics
        self.flux_type = flux_type
        self.airfoil = airfoil
        self.formulation = formulation
        self.optimize = optimize
        self.species = species
        self.equation_classes = [self]


class SimulationMonitor:
    """opensbli/core/diagnostics/simulation_monitors.py:40.  ``arrays`` name the
    probed datasets (``"rho_B0"``, ``"rho"`` or DataObjects), ``probe_locations``
    the (i, j, k) grid index of each.  The HIP back end prints them every
    ``print_frequency`` iterations and after the first, as the generated
    program does, to ``block<b>_<output_file>`` (``output_file=None``: stdout)."""

    def __init__(self, arrays, probe_locations, blocks, print_frequency=100, OPS_V2=True, fp_precision=12,
                 NaN_check=None, output_file="output.log"):
        arrays, probe_locations = list(arrays), list(probe_locations)
        if len(arrays) != len(probe_locations):
            raise ValueError(f"one probed array per probe location: {len(arrays)} arrays for {len(probe_locations)} locations")
        self.blocks = _flatten([blocks])
        nd = self.blocks[0].ndim
        for loc in probe_locations:
            if isinstance(loc, tuple) and len(loc) != nd:
                raise ValueError(f"probe location {loc} has {len(loc)} indices for a {nd}-D block")
        self.arrays = arrays
        self.probe_locations = probe_locations
        self.frequency = int(print_frequency)
        self.fp_precision = int(fp_precision)
        self.NaN_check = NaN_check
        self.output_file = output_file


class TraditionalAlgorithmRK:
    """opensbli/code_generation/algorithm/algorithm.py:334 (single block)."""

    def __init__(self, blocks, simulation_monitor=None):
        blocks = _flatten([blocks])
        if len(blocks) != 1:
            raise NotImplementedError("one block per algorithm")
        self.blocks = blocks
        self.simulation_monitor = simulation_monitor


def _flatten(x) -> list:
    out = []
    for e in x:
        if isinstance(e, (list, tuple)):
            out += _flatten(e)
        else:
            out.append(e)
    return out


__all__ = [n for n in dir() if not n.startswith("_")]
