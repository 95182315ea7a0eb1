"""OpenSBLI's problem-definition surface, targeting the MI355X kernels.

A problem script written for OpenSBLI (apps/*/*.py) runs against this module
with one change: the code-generation target ``OPSC(alg)`` becomes
``HIPC(alg)``.  The objects keep the reference's names and call signatures:

    EinsteinEquation.expand            parsing.py:379
    SimulationEquations, ConstituentRelations   opensbliequations.py:214,444
    SimulationBlock                    block.py:51
    Central, LLFWeno, LLFTeno, StoreSome        scheme.py:63, weno.py:468, teno.py:468, StoreSome.py:10
    RoeAverage, SimpleAverage          averaging.py:31,62
    RungeKutta, RungeKuttaLS           rk_sbli.py:16, rk_LS.py:15
    PeriodicBC, ExtrapolationBC        periodic.py:7, extrapolation.py:8
    ZeroGradientOutletBC, AdiabaticWallBC, IsothermalWallBC, DirichletBC,
    InletPressureExtrapolateBC         zero_gradient_outlet.py:5, adiabatic_wall.py:10,
                                       isothermal_wall.py:11, dirichlet.py:6,
                                       inlet_pressure_extrapolate.py:11
    Carpenter, ReducedAccess           Carpenter_scheme.py:6, reduced_access_scheme.py:6
    MetricsEquation                    metric.py:20 (directions stretched along themselves)
    GridBasedInitialisation            gridbasedinit.py:15
    Initialise_Katzer                  katzer_init.py:183
    TraditionalAlgorithmRK             algorithm.py:336
    WENOFilter, BinomialFilter, SFD    WENO_filter.py:15, binomial_filter.py:15, SFD.py:27
    substitute_simulation_parameters   helperfunctions.py:130

Instead of expanding the Einstein notation symbolically, HIPC recognises the
equation systems the GPU kernels implement (the compressible Euler/Navier-
Stokes equations in skew-symmetric or conservative form, the constituent
relations, the constant or Sutherland-viscosity stress tensor and heat flux of
the reference's apps) and refuses anything else with the reference's error
types.
The initial condition equations are evaluated on the host with SymPy.
"""
import functools
import math
import re

import numpy as np
import sympy
from sympy import Eq, Piecewise, flatten, pprint, symbols  # noqa: F401  (re-exported like opensbli's namespace)
from sympy.parsing.sympy_parser import parse_expr  # noqa: F401

from . import native

# ---------------------------------------------------------------------------
# symbols of the problem script
# ---------------------------------------------------------------------------
_KINDS = {}


def _mk(name, kind):
    name = str(name)
    s = sympy.Symbol(name)
    _KINDS[name] = kind
    return s


class _GridSymbol(sympy.Symbol):
    """A kernel-local variable: a symbol distinct from the dataset of the same
    name (Eq(DataObject(rho), GridVariable(rho)) is an assignment)."""


def GridVariable(name):
    """Kernel-local variable of an initial condition (grid.py:214)."""
    name = str(name)
    _KINDS.setdefault(name, "grid")
    return _GridSymbol(name)


def DataObject(name):
    """A dataset of the block (opensbliobjects.py:398)."""
    return _mk(name, "data")


def ConstantObject(name, **kwargs):
    """A simulation constant (opensbliobjects.py:128)."""
    return _mk(name, "constant")


OpenSBLIEq = Eq


class Double:
    """FP64 datatype (datatypes.py:14)."""


class FloatC:
    """FP32 datatype; the MI355X kernels compute in FP64 only."""


class SimulationDataType:
    dtype = Double

    @classmethod
    def set_datatype(cls, t):
        if t is not Double:
            raise NotImplementedError("the HIPC target computes in FP64 (Double) only")
        cls.dtype = t


# ---------------------------------------------------------------------------
# equations
# ---------------------------------------------------------------------------
class ExpandedEquation:
    """An equation string recorded by EinsteinEquation.expand, with its context."""

    def __init__(self, text, ndim, coordinate_symbol, substitutions, constants):
        self.text = text
        self.ndim = ndim
        self.coordinate_symbol = coordinate_symbol
        self.substitutions = list(substitutions)
        self.constants = list(constants)

    def __repr__(self):
        return "ExpandedEquation(%r)" % self.text


class EinsteinEquation:
    """Records an equation in Einstein notation (parsing.py:379-454).  The
    index expansion is implied by the recognised form."""

    def expand(self, expression, ndim, coordinate_symbol, substitutions=(), constants=(), Metric=None):
        if Metric is not None:
            raise NotImplementedError("metric transformations are not supported by the HIPC target yet")
        return ExpandedEquation(expression, ndim, coordinate_symbol, substitutions, constants)


class _EquationSet:
    def __init__(self):
        self.equations = []

    def add_equations(self, eqns):
        if isinstance(eqns, (list, tuple)):
            self.equations.extend(flatten(list(eqns)))
        else:
            self.equations.append(eqns)

    def write_latex(self, latex):
        return


class SimulationEquations(_EquationSet):
    """The time-advanced equations (opensbliequations.py:214)."""

    def __init__(self, order=None, **kwargs):
        super().__init__()

    def apply_metrics(self, metriceq):
        """Transform the equations with the block's metric (opensbliequations.py:93-110)."""
        if metriceq.stretching_metric is None:
            raise ValueError("generate_transformations must be called before apply_metrics")
        self.metric = metriceq

    @property
    def time_advance_arrays(self):
        """The conservative arrays (opensbliequations.py:426-438)."""
        nd = getattr(self.equations[0], "ndim", None) if self.equations else None
        if nd is None:
            return [DataObject("rho"), DataObject("rhou"), DataObject("rhoE")]
        return [DataObject("rho")] + [DataObject("rhou%d" % d) for d in range(nd)] + [DataObject("rhoE")]


class ConstituentRelations(_EquationSet):
    """Constituent relations (opensbliequations.py:444)."""


class GridBasedInitialisation(_EquationSet):
    """Initial-condition equations evaluated over the grid before the time
    loop (gridbasedinit.py:15-67); here evaluated on the host with SymPy."""


class Initialise_Katzer:
    """Laminar compressible boundary-layer initial state (katzer_init.py:183):
    the similarity solution at (Re, xMach, Tinf) with an adiabatic wall, on the
    wall-normal coordinate given by `coordinate_evaluations`.  One wall
    direction (bl_directions[d][0] True) is implemented."""

    def __init__(self, bl_directions, n_coeffs, Re, xMach, Tinf, coordinate_evaluations=None):
        self.bl_directions = list(bl_directions)
        flags = [bool(b[0]) for b in self.bl_directions]
        if sum(flags) < 1:
            raise ValueError("Provide the directions to apply a boundary layer profile in.")
        if sum(flags) > 1:
            raise NotImplementedError("boundary layers on two walls are not implemented by the HIPC target")
        if n_coeffs < 10:
            raise ValueError("Higher number of polynomial coefficients are required for a good polynomial fit.")
        self.wall_direction = flags.index(True)
        self.Re, self.xMach, self.Tinf = float(Re), float(xMach), float(Tinf)
        self.coordinate_evaluations = list(coordinate_evaluations or [])
        self.equations = self.coordinate_evaluations


# ---------------------------------------------------------------------------
# schemes
# ---------------------------------------------------------------------------
class Central:
    """4th-order central scheme (scheme.py:63)."""

    def __init__(self, order):
        if order != 4:
            raise NotImplementedError("the HIPC target implements the 4th-order central scheme only")
        self.order = order
        self.name = "CentralDerivative"
        self.schemetype = "Spatial"


class StoreSome(Central):
    """Central scheme storing only the listed derivatives (StoreSome.py:10);
    the fused kernels evaluate every derivative in registers, so this is the
    same discretisation."""

    def __init__(self, order, fns):
        super().__init__(order)
        self.fns = fns


class SimpleAverage:
    """Arithmetic interface average (averaging.py:31)."""

    def __init__(self, locations):
        if list(locations) != [0, 1]:
            raise NotImplementedError("averaging locations other than [0, 1]")
        self.kind = "simple"


class RoeAverage:
    """Roe interface average (averaging.py:62)."""

    def __init__(self, locations, physics=None):
        if list(locations) != [0, 1]:
            raise NotImplementedError("averaging locations other than [0, 1]")
        self.kind = "roe"


class LLFWeno:
    """Local Lax-Friedrichs characteristic WENO (weno.py:468)."""

    def __init__(self, order, physics=None, averaging=None, shock_filter=None, formulation="JS"):
        if order != 5:
            raise NotImplementedError("the HIPC target implements WENO5 only")
        if shock_filter is not None:
            raise NotImplementedError("WENO used as a shock filter is not supported by the HIPC target")
        f = formulation.upper()
        if f not in ("Z", "JS"):
            raise NotImplementedError("Only WENO-Z and WENO-JS schemes are implemented.")
        self.conv = "weno_z" if f == "Z" else "weno_js"
        self.averaging = (averaging or SimpleAverage([0, 1])).kind
        self.name = "WenoDerivative"
        self.schemetype = "Spatial"


class LLFTeno:
    """Local Lax-Friedrichs characteristic TENO (teno.py:468)."""

    def __init__(self, order, formulation=None, physics=None, averaging=None, sensor=None, store_sensor=False):
        if order not in (5, 6):
            raise NotImplementedError("Only 5th and 6th order TENO are supported by the HIPC target.")
        if sensor is None and formulation is not None:
            raise ValueError("Storage array for the shock sensor is required.")
        if formulation not in (None, "adaptive"):
            raise NotImplementedError("TENO formulation %r is not implemented by the HIPC target" % formulation)
        self.adaptive = formulation == "adaptive"
        self.sensor_array = sensor
        self.store_sensor = store_sensor
        self.conv = "teno5" if order == 5 else "teno6"
        self.averaging = (averaging or SimpleAverage([0, 1])).kind
        self.name = "TenoDerivative"
        self.schemetype = "Spatial"


class _SensorEquation:
    """The modified Ducros sensor of ShockSensor.ducros_equations as a
    constituent relation (recognised by name)."""

    def __init__(self, name, ndim):
        self.text = "Eq(%s,DUCROS)" % name
        self.ndim = ndim
        self.substitutions = ()


class ShockSensor:
    """Modified Ducros shock sensor (shock_sensors.py:8): theta =
    1/2 (1 - tanh(2.5 (1 + 100 div u))) div^2 / (div^2 + |curl u|^2 + epsilon);
    `epsilon` is a simulation constant."""

    def ducros_equations(self, block, coordinate_symbol, metrics=None, name="theta"):
        if block.ndim < 2:
            raise ValueError("vorticity is not defined for one-dimensional cases.")
        return [_SensorEquation(name, block.ndim)], DataObject(name)


# ---------------------------------------------------------------------------
# filters after the full time step (filters/*.py)
# ---------------------------------------------------------------------------
class _Filter:
    def __init__(self):
        # the script passes equation_classes to block.set_equations, as for
        # the reference's filter objects
        self.equation_classes = [self]


class WENOFilter(_Filter):
    """WENO non-linear filter (WENO_filter.py:15-46): the dissipative part of
    a characteristic WENO5-JS reconstruction (SimpleAverage) against the
    sixth-order central flux, scaled by the non-linear weights' sensor, a
    dissipation sensor ('Ducros' or 'Constant', optionally Yee's low-Mach
    correction) and the constant shock_filter_control; off within 6 points of
    walls.  Applied after every full time step."""

    def __init__(self, block, order, metrics=None, dissipation_sensor="Ducros", Mach_correction=False):
        super().__init__()
        if dissipation_sensor not in ("Ducros", "Constant"):
            raise NotImplementedError("Please enter a valid dissipation sensor option: 'Ducros', or 'Constant'.")
        if order != 5:
            raise NotImplementedError("the HIPC target implements the WENO filter of order 5")
        if block.ndim < 2:
            raise ValueError("vorticity is not defined for one-dimensional cases.")
        block.shock_filter = True
        self.order = order
        self.metrics = metrics
        self.sensor = "ducros" if dissipation_sensor == "Ducros" else "constant"
        self.mach_correction = bool(Mach_correction)


class BinomialFilter(_Filter):
    """Binomial filter of even order (binomial_filter.py:15-34): weights
    binomial(N, i)/2^N per direction, averaged over the directions and blended
    with weight sigma_filt, where grid_condition (a boolean expression of the
    coordinates and grid indices) holds.  Applied last in every time step."""

    def __init__(self, block, order, grid_condition=None, sigma=0.05):
        super().__init__()
        if order % 2 != 0:
            raise ValueError("The filter is only defined for even orders n.")
        if order > 10:
            raise ValueError("Increase the number of halo points in scheme.py for high order filters.")
        self.order = order
        self.grid_condition = grid_condition
        self.sigma = sigma


class SFD(_Filter):
    """Selective frequency damping (SFD.py:27-37): chi_filt, omega_filt;
    the filtered solution starts as the initial state."""

    def __init__(self, block, chifilt=0.1, omegafilt=1.0 / 0.75):
        super().__init__()
        self.chi = chifilt
        self.omega = omegafilt


class RungeKutta:
    """3-register RK3 (rk_sbli.py:16)."""

    def __init__(self, order, constant_dt=None):
        if order != 3:
            raise NotImplementedError("This Runge-Kutta scheme is only defined for 3rd order. "
                                      "For 4th order please use the RungeKuttaLS class instead.")
        self.rk = "rk3"
        self.name = "RungeKutta"
        self.schemetype = "Temporal"


class RungeKuttaLS:
    """Low-storage Williamson RK (rk_LS.py:15)."""

    def __init__(self, order, formulation=None, constant_dt=None):
        if order == 3:
            self.rk = "ls3_ssp" if formulation == "SSP" else "ls3"
        elif order == 4:
            if formulation == "SSP":
                raise ValueError("RK-SSP is only available for the 3rd order scheme.")
            self.rk = "ls4"
        else:
            raise NotImplementedError("Only 3rd and 4th order RK schemes are currently implemented.")
        self.name = "RungeKutta"
        self.schemetype = "Temporal"


# ---------------------------------------------------------------------------
# boundary conditions
# ---------------------------------------------------------------------------
class PeriodicBC:
    """Periodic exchange (periodic.py:7)."""

    def __init__(self, direction, side, plane=True):
        self.direction, self.side, self.kind = direction, side, "periodic"


class Carpenter:
    """Carpenter's 4th-order one-sided closure of the central stencils
    (Carpenter_scheme.py:6), the default of every stencil-modifying BC."""
    name = "carpenter"


class ReducedAccess:
    """The (-4, 4)-access closure (reduced_access_scheme.py:6)."""
    name = "reduced_access"


def _closure(scheme):
    if scheme is None:
        return "carpenter"
    if isinstance(scheme, (Carpenter, ReducedAccess)) or scheme in (Carpenter, ReducedAccess):
        return scheme.name
    raise NotImplementedError("boundary scheme %r is not implemented by the HIPC target" % (scheme,))


class _BC:
    kind = None
    equations = None

    def __init__(self, direction, side, scheme=None, plane=True):
        if not plane:
            raise NotImplementedError("split boundary conditions are not implemented by the HIPC target")
        self.direction, self.side = direction, side
        self.scheme = _closure(scheme)


class ExtrapolationBC(_BC):
    """Extrapolation boundary (extrapolation.py:8): order 0 or 1."""

    def __init__(self, direction, side, order, scheme=None, plane=True):
        super().__init__(direction, side, scheme, plane)
        if order not in (0, 1):
            raise ValueError("Only zeroth order and linear extrapolation currently implemented.")
        self.order = order
        self.kind = "extrap0" if order == 0 else "extrap1"


class ZeroGradientOutletBC(_BC):
    """q[boundary] = q[boundary - 1], halos mirrored (zero_gradient_outlet.py:5);
    leaves the central stencils unmodified."""
    kind = "zero_gradient"


class SymmetryBC(_BC):
    """Symmetry plane (symmetry.py:6): halos mirrored with the normal momentum
    negated; not a ModifyCentralDerivative, so the central stencils read the
    halos (a `scheme` argument is accepted and unused, as in the reference)."""
    kind = "symmetry"


class AdiabaticWallBC(_BC):
    """No-slip wall with dT/dn = 0 (adiabatic_wall.py:10)."""
    kind = "adiabatic_wall"


class IsothermalWallBC(_BC):
    """No-slip wall at T = Twall (isothermal_wall.py:11).  `equations` must set
    rhoE = rho*Twall/(gama*(gama-1)*Minf**2), the form of the reference's apps."""
    kind = "isothermal_wall"

    def __init__(self, direction, side, equations, scheme=None, plane=True):
        super().__init__(direction, side, scheme, plane)
        self.equations = list(flatten([equations]))


class DirichletBC(_BC):
    """Fixed values on the boundary plane and its halos (dirichlet.py:6).  The
    equations may depend on the coordinates, grid indexes and constants."""
    kind = "dirichlet"

    def __init__(self, direction, side, equations, scheme=None, plane=True):
        super().__init__(direction, side, scheme, plane)
        self.equations = list(flatten([equations]))


class InletPressureExtrapolateBC(_BC):
    """Inlet with pressure extrapolation when subsonic (inlet_pressure_extrapolate.py:11)."""
    kind = "inlet_pextrap"

    def __init__(self, direction, side, scheme=None, plane=True):
        super().__init__(direction, side, scheme, plane)
        if side != 0:
            raise ValueError("Only implemented this BC for inlet side 0.")


# ---------------------------------------------------------------------------
# metric
# ---------------------------------------------------------------------------
class MetricsEquation:
    """Coordinate transformation of the derivatives (metric.py:20).  The HIPC
    target implements directions stretched along themselves, parameters
    (stretched, False): the metric is D_dd = 1/CD(x_d, xi_d) with
    SD_ddd = CD(D_dd, xi_d), computed from the block's coordinates; and the
    general curvilinear metric, every direction curvilinear: D = adj(J)/detJ
    with J_ab = CD(x_a, xi_b) (metric.py:72-147), used by equations written in
    strong conservation form in xi (apps/euler_wave_curvilinear/euler_wave.py)
    on periodic blocks."""

    def __init__(self, **kwargs):
        self.stretching_metric = None
        self.curvilinear_metric = None
        self.ndim = None
        self.metric_subs = {}

    def generate_transformations(self, ndim, coordinate_symbol, parameters, max_order):
        if len(flatten(list(parameters))) != ndim * 2:
            raise ValueError("The parameters for stretching provided should match the number of dimensions")
        self.ndim = ndim
        self.stretching_metric = [bool(p[0]) for p in parameters]
        self.curvilinear_metric = [bool(p[1]) for p in parameters]
        if any(self.curvilinear_metric) and not all(self.curvilinear_metric):
            raise NotImplementedError("the HIPC target implements curvilinear metrics in every direction or in none")
        self.coordinate_symbol = coordinate_symbol
        return


# ---------------------------------------------------------------------------
# I/O and documentation hooks of the reference, accepted and recorded
# ---------------------------------------------------------------------------
class BeforeSimulationStarts:
    """Placement before the time loop (common.py:5)."""


class AfterSimulationEnds:
    """Placement after the time loop (common.py:11)."""


class InTheSimulation:
    """Placement inside the time loop, every `frequency` iterations (common.py:21)."""

    def __init__(self, frequency=False):
        self.frequency = frequency


class iohdf5:
    """HDF5 input/output of block arrays (io_hdf5.py:19-138), in the OPS file
    layout (opensbli_amd.hdf5).

    iotype "write" (the default): after the time loop the arrays go to
    "opensbli_output.h5" (or the `name` keyword); with save_every, also after
    every iteration k with k % save_every == 0, to "opensbli_output_%06d.h5" % k.
    iotype "read": before the loop the arrays are read from `filename`
    (default "data.h5"): conservative arrays the initial condition does not set
    become the initial state (a restart), other arrays are available to the
    initial-condition equations (the KH app's random_nums).  Keyword values are
    lower-cased, as the reference does (io_hdf5.py:26-29)."""

    def __init__(self, arrays=None, save_every=None, **kwargs):
        if kwargs:
            self.kwargs = {key.lower(): str(val).lower() for key, val in kwargs.items()}
        else:
            self.kwargs = {"iotype": "write"}
        self.save_every = save_every
        self.algorithm_place = []
        if save_every:
            self.algorithm_place.append(InTheSimulation(save_every))
        if self.kwargs["iotype"] == "write":
            self.algorithm_place.append(AfterSimulationEnds())
        elif self.kwargs["iotype"] == "read":
            self.algorithm_place = [BeforeSimulationStarts()]
        else:
            raise ValueError("")
        self.arrays = []
        if arrays:
            self.add_arrays(arrays)

    def add_arrays(self, arrays):
        self.arrays.extend(flatten(list(arrays)) if isinstance(arrays, (list, tuple)) else [arrays])

    @property
    def iotype(self):
        return self.kwargs["iotype"]

    def file_name(self, iteration=None):
        """Output file of a write, after the loop (iteration None) or after
        `iteration` iterations inside it (io_hdf5.py:99-118)."""
        if "name" in self.kwargs:
            name = self.kwargs["name"]
            if ".h5" in name:
                fname = name
            elif "." in name:
                raise ValueError("")
            else:
                fname = name + ".h5"
            if iteration is not None:
                raise ValueError("dynamic fname not allowed ")
            return fname
        if iteration is None:
            return "opensbli_output.h5"
        return "opensbli_output_%06d.h5" % iteration

    def read_file_name(self):
        return self.kwargs.get("filename", "data.h5")

    def check_datasets(self, defined):
        """Refuse arrays the simulation does not define (io_hdf5.py:55-62)."""
        missing = [x for x in _array_names(self.arrays) if x not in defined]
        if missing:
            raise ValueError("The dataset(s): '%s' added to the HDF5 class are not defined in the simulation code. "
                             "Please check the HDF5 add_arrays input in the problem script." % str(", ".join(missing)))


def _array_names(arrays, ndim=None):
    """Array names of an iohdf5 list: symbols or strings, a trailing _B<n>
    dropped, "rhou" expanded to its components when ndim is given."""
    out = []
    for a in arrays:
        nm = re.sub(r"_B\d+$", "", str(a))
        if nm == "rhou" and ndim:
            out.extend("rhou%d" % d for d in range(ndim))
        else:
            out.append(nm)
    return out


def output_hdf5(array, array_name, halos, npoints, block, **kwargs):
    """Write arrays (with their halo layers) to an HDF5 file a simulation reads
    with iohdf5(iotype="Read") (helperfunctions.py:103-127), default "data.h5".

    halos: one (low, high) pair per direction, e.g. [(-5, 5), (-5, 5)]; npoints
    the interior points per direction; each array's axes in reverse direction
    order (x last), of extent npoints + halos.  On a 2-D block the argument
    checks are the reference's (helperfunctions.py:80-88)."""
    from . import hdf5
    if not isinstance(array, list):
        array = [array]
    if not isinstance(array_name, list):
        array_name = [array_name]
    assert len(array) == len(array_name)
    nd = block.ndim
    if nd == 2 or len(halos) == 2:
        if len(halos) != 2:
            raise ValueError("Two halos should be provided for each dimension.")
        for h in halos:
            if len(h) != nd:
                raise ValueError("halos provided for hdf5 output should be of size %d" % nd)
    elif len(halos) != nd or any(len(h) != 2 for h in halos):
        raise ValueError("halos provided for hdf5 output should be one (low, high) pair per dimension (%d)" % nd)
    d_m = [int(halos[d][0]) for d in range(nd)]
    d_p = [int(halos[d][1]) for d in range(nd)]
    fname = kwargs.get("filename", "data.h5")
    hdf5.fetch_block_hdf5_file(fname, block.blockname, block.blocknumber, nd, truncate=True)
    for ar, name in zip(array, array_name):
        ar = np.asarray(ar, dtype=np.float64)
        if ar.ndim != nd:
            raise ValueError("array %s has %d dimensions, the block %d" % (name, ar.ndim, nd))
        size = [ar.shape[nd - 1 - d] - abs(d_m[d]) - abs(d_p[d]) for d in range(nd)]
        if list(size) != [int(v) for v in npoints]:
            raise ValueError("array %s: extent %s minus halos %s is not npoints %s"
                             % (name, ar.shape, halos, list(npoints)))
        hdf5.fetch_dat_hdf5_file(fname, block.blockname, block.blocknumber, hdf5.dataset_name(name, block.blocknumber),
                                 ar, [-abs(v) for v in d_m], [abs(v) for v in d_p])
    return


class LatexWriter:
    """Equation documentation writer (latex.py); writes nothing here."""

    def open(self, *a, **k):
        return

    def write_string(self, *a, **k):
        return

    def write_expression(self, *a, **k):
        return

    def close(self):
        return


# ---------------------------------------------------------------------------
# block and algorithm
# ---------------------------------------------------------------------------
class _Deltas:
    def __init__(self, ndim, blocknumber):
        self._s = [sympy.Symbol("Delta%dblock%d" % (d, blocknumber)) for d in range(ndim)]

    def __getitem__(self, i):
        return self._s[i]


class SimulationBlock:
    """A structured block (block.py:51)."""

    def __init__(self, ndim, block_number=0):
        self.ndim = ndim
        self.blocknumber = block_number
        self.blockname = "opensbliblock%02d" % block_number
        self.deltas = _Deltas(ndim, block_number)
        self.grid_indexes = [sympy.Symbol("idx%d" % d) for d in range(ndim)]
        self.boundaries = None
        self.equation_classes = []
        self.discretisation_schemes = {}
        self.io = []
        self.discretised = False
        self.shock_filter = False

    def set_block_boundaries(self, bclist):
        flat = flatten(list(bclist))
        table = [[None, None] for _ in range(self.ndim)]
        for bc in flat:
            table[bc.direction][bc.side] = bc
        for d in range(self.ndim):
            for s in range(2):
                if table[d][s] is None:
                    raise ValueError("boundary condition missing for direction %d side %d" % (d, s))
            if (table[d][0].kind == "periodic") != (table[d][1].kind == "periodic"):
                raise ValueError("direction %d: periodic boundaries must be set on both sides" % d)
        self.boundaries = table

    def set_equations(self, list_of_equations):
        self.equation_classes = list(flatten(list(list_of_equations)))

    def set_discretisation_schemes(self, schemes):
        self.discretisation_schemes = dict(schemes)

    def setio(self, io):
        self.io = list(io) if isinstance(io, (list, tuple)) else [io]

    def discretise(self):
        if self.boundaries is None:
            raise ValueError("set_block_boundaries must be called before discretise")
        if not self.discretisation_schemes:
            raise ValueError("set_discretisation_schemes must be called before discretise")
        self.discretised = True


class SimulationMonitor:
    """Probes of conservative variables at grid points, printed during the
    run, with an optional NaN check (core/diagnostics/simulation_monitors.py:16-44).
    After iteration 1 and every print_frequency-th iteration: the NaN check of
    the first monitored array (NaNcheck=True; ops_NaNcheck stops the run, here
    FloatingPointError), the header 'Iteration, Time, name(i, j, k), ...' at
    the first print, then 'iteration, time, values...' with fp_precision
    decimals, to stdout or appended to output_file (simulation_monitors.py:
    112-160, the condition at algorithm.py:433-437)."""

    def __init__(self, arrays, probe_locations, block, print_frequency=250, fp_precision=10, NaNcheck=True,
                 output_file=None):
        if len(arrays) != len(probe_locations):
            raise ValueError("The number of arrays must equal the number of probe locations.")
        self.arrays = [str(a) for a in arrays]
        self.probe_locations = [tuple(int(i) for i in loc) if isinstance(loc, (tuple, list)) else (int(loc),)
                                for loc in probe_locations]
        self.block = block
        self.ndim = block.ndim
        self.frequency = int(print_frequency)
        self.fp_precision = int(fp_precision)
        self.NaNcheck = NaNcheck
        self.output_file = output_file
        names = _state_names(self.ndim)
        for a, loc in zip(self.arrays, self.probe_locations):
            if a not in names:
                raise NotImplementedError("the HIPC target monitors the conservative arrays %s, not %s" % (names, a))
            if len(loc) != self.ndim:
                raise ValueError("probe location %s of %s has %d indices on a %d-D block"
                                 % (loc, a, len(loc), self.ndim))

    def header(self):
        return ", ".join(["Iteration", "Time"] + ["%s%s" % (a, loc if len(loc) > 1 else "(%d)" % loc[0])
                                                   for a, loc in zip(self.arrays, self.probe_locations)])

    def line(self, iteration, time, values):
        f = "%%.%df" % self.fp_precision
        return ", ".join(["%d" % iteration, f % time] + [f % v for v in values])


def _state_names(nd):
    return ["rho"] + ["rhou%d" % d for d in range(nd)] + ["rhoE"]


class TraditionalAlgorithmRK:
    """The RK time-loop algorithm of a block (algorithm.py:336)."""

    def __init__(self, blocks, simulation_monitor=None, dtype=None):
        if isinstance(blocks, (list, tuple)):
            raise NotImplementedError("Multi-Block algorithm is not implemented")
        if not blocks.discretised:
            raise ValueError("discretise the block before building the algorithm")
        self.block = blocks
        self.simulation_monitor = simulation_monitor


# ---------------------------------------------------------------------------
# recognition of the equation system
# ---------------------------------------------------------------------------
def _norm(text):
    t = re.sub(r"\s+", "", text)
    t = re.sub(r",\*\*\{['\"]scheme['\"]:['\"](\w+)['\"]\}", r"@\1", t)
    return t


# operators of the equation language: kept as undefined functions, so that
# only the algebra around and inside them is normalised
_OPERATORS = ("Der", "Skew", "Conservative", "KD", "Dot", "Eq")


@functools.lru_cache(maxsize=None)
def _canon(text):
    """Canonical form of an expression (or Eq) of the equation language:
    sympy's ordering of sums and products with every product expanded, also
    inside the operators' arguments, so that algebraically identical forms
    (reordered factors or terms, a factor distributed over a sum) compare
    equal.  Every name is a plain symbol or an undefined function (none takes
    sympy's meaning, such as E or beta); the scheme keyword of Conservative
    becomes a symbol argument.  Text that does not parse is compared as
    written."""
    t = _norm(text)
    t = re.sub(r"@(\w+)", r",scheme_\1", t)
    names = set(re.findall(r"[A-Za-z_]\w*", t))
    local = {n: (sympy.Function(n) if n in _OPERATORS else sympy.Symbol(n)) for n in names}
    try:
        e = parse_expr(t, local_dict=local)
    except Exception:  # noqa: BLE001 (anything sympy cannot read is compared as text)
        return t
    return sympy.srepr(sympy.expand(e, deep=True))


def _canon_term(sign, text):
    """(sign, canonical form) of a top-level term, a leading minus sign of the
    term itself moved into `sign`."""
    t = _norm(text)
    t = re.sub(r"@(\w+)", r",scheme_\1", t)
    names = set(re.findall(r"[A-Za-z_]\w*", t))
    local = {n: (sympy.Function(n) if n in _OPERATORS else sympy.Symbol(n)) for n in names}
    try:
        e = sympy.expand(parse_expr(t, local_dict=local), deep=True)
    except Exception:  # noqa: BLE001
        return sign, _norm(text)
    if e.could_extract_minus_sign():
        sign, e = -sign, -e
    return sign, sympy.srepr(e)


def _terms(pairs):
    """Canonical (sign, term) pairs of a set of (sign, text) pairs."""
    return {_canon_term(sg, tx) for sg, tx in pairs}


def _split_terms(rhs):
    """Top-level signed terms of an expression string."""
    terms, depth, cur, sign = [], 0, "", 1
    for ch in rhs:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if depth == 0 and ch in "+-":
            if cur:
                terms.append((sign, cur))
            sign = 1 if ch == "+" else -1
            cur = ""
            continue
        cur += ch
    if cur:
        terms.append((sign, cur))
    return terms


def _parse_eq(text):
    t = _norm(text)
    m = re.fullmatch(r"Eq\((.*)\)", t)
    if not m:
        raise ValueError("not an equation: %r" % text)
    body = m.group(1)
    depth = 0
    for i, ch in enumerate(body):
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        elif ch == "," and depth == 0:
            return body[:i], body[i + 1:]
    raise ValueError("not an equation: %r" % text)


SKEW = {
    "rho": {(-1, "Skew(rho*u_j,x_j)")},
    "rhou_i": {(-1, "Skew(rhou_i*u_j,x_j)"), (-1, "Der(p,x_i)")},
    "rhoE": {(-1, "Skew(rhoE*u_j,x_j)"), (-1, "Conservative(p*u_j,x_j)")},
}


def _cons(tag):
    return {
        "rho": {(-1, "Conservative(rhou_j,x_j%s)" % tag)},
        "rhou_i": {(-1, "Conservative(rhou_i*u_j+KD(_i,_j)*p,x_j%s)" % tag)},
        "rhoE": {(-1, "Conservative((p+rhoE)*u_j,x_j%s)" % tag)},
    }


# strong conservation form in curvilinear coordinates xi
# (apps/euler_wave_curvilinear/euler_wave.py:17-25)
def _curv(tag):
    return {
        "rho": {(-1, "Conservative(detJ*rho*U_j,xi_j%s)/detJ" % tag)},
        "rhou_i": {(-1, "Conservative(detJ*(rhou_i*U_j+p*D_j_i),xi_j%s)/detJ" % tag)},
        "rhoE": {(-1, "Conservative(detJ*(p+rhoE)*U_j,xi_j%s)/detJ" % tag)},
    }


VISCOUS = {"rhou_i": {(1, "Der(tau_i_j,x_j)")}, "rhoE": {(1, "Der(u_i*tau_i_j,x_j)")}}
STRESS = "Eq(tau_i_j,(1.0/Re)*(Der(u_i,x_j)+Der(u_j,x_i)-(2/3)*KD(_i,_j)*Der(u_k,x_k)))"
HEAT_POS = "Eq(q_j,(1.0/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
HEAT_NEG = "Eq(q_j,(-1.0/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
# Sutherland viscosity as a constituent relation (katzer_SBLI.py:12-26)
STRESS_MU = "Eq(tau_i_j,(mu/Re)*(Der(u_i,x_j)+Der(u_j,x_i)-(2/3)*KD(_i,_j)*Der(u_k,x_k)))"
HEAT_MU_POS = "Eq(q_j,(mu/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
HEAT_MU_NEG = "Eq(q_j,(-mu/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
SUTHERLAND = "(T**(1.5)*(1.0+SuthT/RefT)/(T+SuthT/RefT))"
CONSTITUENT = {
    "u_i": "rhou_i/rho",
    "p": "(gama-1)*(rhoE-rho*(1/2)*(KD(_i,_j)*u_i*u_j))",
    "a": "(gama*p/rho)**0.5",
    "T": "p*gama*Minf*Minf/(rho)",
    "U_i": "D_i_j*u_j",
}


def _recognise(block):
    sim = [e for e in block.equation_classes if isinstance(e, SimulationEquations)]
    crs = [e for e in block.equation_classes if isinstance(e, ConstituentRelations)]
    if len(sim) != 1:
        raise ValueError("exactly one SimulationEquations object is required on the block")
    eqs = sim[0].equations
    ndim = block.ndim
    terms = {}
    subs = set()
    for e in eqs:
        if e.ndim != ndim:
            raise ValueError("equation expanded for %d dimensions on a %d-D block" % (e.ndim, ndim))
        lhs, rhs = _parse_eq(e.text)
        m = re.fullmatch(r"Der\((\w+),t\)", lhs)
        if not m:
            raise NotImplementedError("left-hand side %r is not a time derivative" % lhs)
        terms[m.group(1)] = _terms(_split_terms(rhs))
        subs.update(_canon(s) for s in e.substitutions)
    if set(terms) != {"rho", "rhou_i", "rhoE"}:
        raise NotImplementedError("the HIPC target solves the mass, momentum and energy equations; got %s"
                                  % sorted(terms))
    # convective family
    schemes = block.discretisation_schemes
    spatial = [s for s in schemes.values() if getattr(s, "schemetype", "") == "Spatial"]
    temporal = [s for s in schemes.values() if getattr(s, "schemetype", "") == "Temporal"]
    if len(temporal) != 1:
        raise ValueError("Found more than one temporal scheme on the block" if temporal else
                         "no temporal scheme on the block")
    # a constant body force (apps/channel_flow): -KD(_i,_j)*c_j and -Dot(c_j, u_j)
    fm, fe = _canon_term(-1, "KD(_i,_j)*c_j"), _canon_term(-1, "Dot(c_j,u_j)")
    force = fm in terms["rhou_i"] and fe in terms["rhoE"]
    if force:
        terms["rhou_i"].discard(fm)
        terms["rhoE"].discard(fe)
    elif fm in terms["rhou_i"] or fe in terms["rhoE"]:
        raise NotImplementedError("a body force needs both -KD(_i,_j)*c_j and -Dot(c_j, u_j)")
    viscous_terms = {k: _terms(v) for k, v in VISCOUS.items()}
    heat_terms = {sg: _canon_term(sg, "Der(q_j,x_j)") for sg in (1, -1)}
    has_visc = any(t in terms[k] for k, v in viscous_terms.items() for t in v)
    heat = None
    if has_visc:
        for sign in (1, -1):
            if heat_terms[sign] in terms["rhoE"]:
                heat = sign
    conv = None
    averaging = "roe"
    for name, want in [("skew", SKEW), ("weno", _cons("@Weno")), ("teno", _cons("@Teno")), ("central", _cons("")),
                       ("weno_curv", _curv("@Weno")), ("teno_curv", _curv("@Teno"))]:
        ok = True
        for var in ("rho", "rhou_i", "rhoE"):
            got = set(terms[var])
            if has_visc:
                got -= viscous_terms.get(var, set())
                if var == "rhoE" and heat is not None:
                    got.discard(heat_terms[heat])
            if got != _terms(want[var]):
                ok = False
        if ok:
            conv = name
            break
    if conv is None:
        raise NotImplementedError("equation system not recognised by the HIPC target: %s" % terms)
    curvilinear = conv.endswith("_curv")
    if curvilinear:
        conv = conv[:-5]
        if has_visc:
            raise NotImplementedError("viscous terms on curvilinear blocks are not implemented by the HIPC target")
        metrics = [e for e in block.equation_classes if isinstance(e, MetricsEquation)]
        if not metrics or not metrics[0].curvilinear_metric or not all(metrics[0].curvilinear_metric):
            raise ValueError("equations in curvilinear coordinates need a MetricsEquation with curvilinear "
                             "directions on the block")
        if metrics[0].ndim != ndim:
            raise ValueError("the metric was generated for %d dimensions on a %d-D block" % (metrics[0].ndim, ndim))
        if not [s for s in spatial if isinstance(s, Central)]:
            raise ValueError("A Central scheme should be instantiated for metrics.")
    if conv in ("skew", "central"):
        cen = [s for s in spatial if isinstance(s, Central)]
        if not cen:
            raise ValueError("a Central(4) scheme is required for the central derivatives")
        conv = "central_skew" if conv == "skew" else "central_cons"
    else:
        cls = LLFWeno if conv == "weno" else LLFTeno
        sc = [s for s in spatial if isinstance(s, cls)]
        if not sc:
            raise ValueError("the equations ask for %s but no such scheme is on the block" % cls.__name__)
        averaging = sc[0].averaging
        conv = sc[0].conv
        if has_visc and not [s for s in spatial if isinstance(s, Central)]:
            raise ValueError("viscous terms need a Central(4) scheme on the block")
    viscous = "none"
    if has_visc:
        for var, want in viscous_terms.items():
            if not want <= terms[var]:
                raise NotImplementedError("partial viscous terms are not supported")
        if _canon(STRESS) in subs:
            heat_ok = (heat == 1 and _canon(HEAT_POS) in subs) or (heat == -1 and _canon(HEAT_NEG) in subs)
            viscous = "const"
        elif _canon(STRESS_MU) in subs:
            heat_ok = (heat == 1 and _canon(HEAT_MU_POS) in subs) or (heat == -1 and _canon(HEAT_MU_NEG) in subs)
            viscous = "sutherland"
        else:
            raise NotImplementedError("only the stress tensors (1.0/Re)*(...) and (mu/Re)*(...) are supported")
        if not heat_ok:
            raise NotImplementedError("heat flux form not supported")
    # constituent relations
    have = {}
    for c in crs:
        for e in c.equations:
            lhs, rhs = _parse_eq(e.text)
            have[lhs] = rhs
    # mu/Re with mu a declared constant rather than a constituent relation:
    # constant viscosity at the effective Reynolds number Re/mu
    mu_constant = viscous == "sutherland" and "mu" not in have and any("mu" in e.constants for e in eqs)
    if mu_constant:
        viscous = "const"
    need = ["u_i", "p"] + (["a"] if conv not in ("central_skew", "central_cons") else []) + \
        (["T"] if viscous != "none" else []) + (["mu"] if viscous == "sutherland" else []) + \
        (["U_i"] if curvilinear else [])
    for k in need:
        if k not in have:
            raise ValueError("Constituent relation is not found for %s" % k)
        want = SUTHERLAND if k == "mu" else CONSTITUENT[k]
        if _canon(have[k]) != _canon(want):
            raise NotImplementedError("constituent relation for %s not supported: %s" % (k, have[k]))
    sensor = 0
    tenos = [s for s in spatial if isinstance(s, LLFTeno)]
    if tenos and tenos[0].adaptive and conv in ("teno5", "teno6"):
        name = str(tenos[0].sensor_array)
        if have.get(name) != "DUCROS":
            raise ValueError("the adaptive TENO needs the shock sensor %s among the constituent relations" % name)
        sensor = 1
    stretch = [0, 0, 0]
    metric = getattr(sim[0], "metric", None)
    if metric is not None:
        if metric.ndim != ndim:
            raise ValueError("the metric was generated for %d dimensions on a %d-D block" % (metric.ndim, ndim))
        if not [s for s in spatial if isinstance(s, Central)]:
            raise ValueError("A Central scheme should be instantiated for metrics.")
        if any(metric.curvilinear_metric):
            raise NotImplementedError("equations in x transformed by a curvilinear metric (apply_metrics) are not "
                                      "implemented by the HIPC target; write them in strong conservation form in "
                                      "xi as apps/euler_wave_curvilinear does")
        for d in range(ndim):
            stretch[d] = int(metric.stretching_metric[d])
    out = dict(conv=conv, averaging=averaging, viscous=viscous, rk=temporal[0].rk, stretch=stretch, sensor=sensor)
    if mu_constant:
        out["mu_constant"] = 1
    if force:
        out["force"] = 1
    if curvilinear:
        out["curvilinear"] = 1
    return out


# ---------------------------------------------------------------------------
# the target
# ---------------------------------------------------------------------------
_CURRENT = {}


def _lookup(env, sym):
    """The value of a symbol of an initial-condition expression: the symbol
    itself, else a kernel-local variable of the same name (a bare name in the
    generated C refers to the local variable; datasets print as name_B0)."""
    if sym in env:
        return env[sym]
    if not isinstance(sym, _GridSymbol):
        return env.get(_GridSymbol(sym.name))
    return None


def _adaptive_ct(theta, a1, a2):
    """TENO_CT = 1/10^floor(a1 - a2 (1 - (1 - theta)^4 (1 + 4 theta))) (teno.py:430-438)."""
    beta = a1 - a2 * (1.0 - (1.0 - theta) ** 4 * (1.0 + 4.0 * theta))
    return 1.0 / 10.0 ** np.floor(beta)


def _metric_names(nd):
    """Datasets of the curvilinear metric, in the order of osbl_get_metric_field."""
    return ["D%d%d" % (a, b) for a in range(nd) for b in range(nd)] + ["detJ"]


def _c_eval(expr, env):
    """Evaluate a C constant expression of substitute_simulation_parameters."""
    names = dict(env)
    names.update(M_PI=math.pi, ceil=math.ceil, floor=math.floor, sqrt=math.sqrt, pow=math.pow, exp=math.exp,
                 sin=math.sin, cos=math.cos, tanh=math.tanh, fabs=abs)
    return eval(expr, {"__builtins__": {}}, names)


class HIPC:
    """Code-generation target replacing OPSC (opsc.py:250): builds the
    configuration of libosbl_hip.so for the algorithm instead of writing C.

    After substitute_simulation_parameters(constants, values), run() evaluates
    the initial condition, executes niter iterations on the GPU and returns the
    final conservative arrays."""

    def __init__(self, algorithm, operation_count=False, OPS_diagnostics=1, simulation_name="opensbli"):
        self.alg = algorithm
        self.block = algorithm.block
        self.scheme = _recognise(self.block)
        self.values = {}
        self.simulation_name = simulation_name
        self.diagnostics = OPS_diagnostics
        _CURRENT[simulation_name] = self

    # ----------------------------------------------------------------------
    def set_values(self, constants, values):
        pending = dict(zip(constants, values))
        for _ in range(len(pending) + 1):
            progress = False
            for k in list(pending):
                try:
                    self.values[k] = _c_eval(str(pending[k]), self.values)
                    del pending[k]
                    progress = True
                except NameError:
                    continue
            if not pending or not progress:
                break
        if pending:
            raise ValueError("cannot evaluate simulation parameters: %s" % pending)

    def problem(self):
        b = self.block
        v = self.values
        nd = b.ndim
        try:
            n = [int(v["block%dnp%d" % (b.blocknumber, d)]) for d in range(nd)]
            delta = [float(v["Delta%dblock%d" % (d, b.blocknumber)]) for d in range(nd)]
            dt = float(v["dt"])
        except KeyError as e:
            raise ValueError("simulation parameter %s has no value" % e)
        p = dict(ndim=nd, n=n, delta=delta, dt=dt, gama=float(v.get("gama", 1.4)),
                 bc=[[b.boundaries[d][s].kind for s in range(2)] for d in range(nd)],
                 bscheme=[[getattr(b.boundaries[d][s], "scheme", "carpenter") for s in range(2)] for d in range(nd)],
                 **self.scheme)
        for k in ("Minf", "Re", "Pr", "eps"):
            if k in v:
                p[k] = float(v[k])
        if p.pop("force", None):
            missing = ["c%d" % d for d in range(nd) if "c%d" % d not in v]
            if missing:
                raise ValueError("simulation parameters %s need values (the body force c_j)" % missing)
            p["body_force"] = [float(v["c%d" % d]) for d in range(nd)]
        if p.pop("mu_constant", None):
            if "mu" not in v or "Re" not in v:
                raise ValueError("simulation parameters mu and Re need values")
            p["Re"] = float(v["Re"]) / float(v["mu"])
        if "TENO_CT" in v:
            p["teno_ct"] = float(v["TENO_CT"])
        if p["viscous"] == "sutherland":
            for k in ("SuthT", "RefT"):
                if k not in v:
                    raise ValueError("simulation parameter %s has no value" % k)
                p[k] = float(v[k])
        if p.get("sensor"):
            for k, name in (("teno_a1", "teno_a1"), ("teno_a2", "teno_a2"), ("sensor_eps", "epsilon")):
                if name not in v:
                    raise ValueError("simulation parameter %s has no value" % name)
                p[k] = float(v[name])
        filters = {}
        for f in self.block.equation_classes:
            if isinstance(f, WENOFilter):
                if float(v.get("sensor_theta", 4.0)) != 4.0:
                    raise NotImplementedError("the HIPC target implements sensor_theta = 4")
                filters["weno"] = {"sensor": f.sensor, "mach_correction": f.mach_correction,
                                   "control": float(v.get("shock_filter_control", 1.0))}
                if f.sensor == "ducros":
                    p["sensor_eps"] = float(v.get("epsilon", 1.0e-12))
            elif isinstance(f, SFD):
                filters["sfd"] = {"chi": float(v.get("chi_filt", f.chi)), "omega": float(v.get("omega_filt", f.omega))}
            elif isinstance(f, BinomialFilter):
                filters["binomial"] = {"order": f.order, "sigma": float(v.get("sigma_filt", f.sigma))}
        if filters:
            p["filters"] = filters
        if any(p["bc"][d][s] == "isothermal_wall" for d in range(nd) for s in range(2)):
            if "Twall" not in v:
                raise ValueError("simulation parameter Twall has no value")
            p["Twall"] = float(v["Twall"])
        return p

    # ----------------------------------------------------------------------
    def _env(self, problem, d_m=None, d_p=None):
        """Symbols of the initial-condition equations over the interior grid,
        or over the grid widened by halo depths d_m (<= 0) and d_p per
        direction: constants, grid indices and the arrays read from HDF5."""
        env = {sympy.Symbol(k): val for k, val in self.values.items()}
        nd = problem["ndim"]
        d_m = d_m or [0] * nd
        d_p = d_p or [0] * nd
        grids = np.meshgrid(*[np.arange(d_m[d], problem["n"][d] + d_p[d]) for d in reversed(range(nd))],
                            indexing="ij")
        for d in range(nd):
            env[sympy.Symbol("idx%d" % d)] = grids[nd - 1 - d].astype(np.float64)
        env[sympy.pi] = math.pi
        for nm, (a, info) in getattr(self, "_inputs", {}).items():
            sl = []
            for d in reversed(range(nd)):
                lo = d_m[d] - info["d_m"][d]
                hi = lo + problem["n"][d] - d_m[d] + d_p[d]
                if lo < 0 or hi > a.shape[nd - 1 - d]:
                    raise ValueError("%s is read with halos %s/%s in direction %d; %s/%s are needed"
                                     % (nm, info["d_m"][d], info["d_p"][d], d, d_m[d], d_p[d]))
                sl.append(slice(lo, hi))
            env[sympy.Symbol(nm)] = a[tuple(sl)]
        return env, grids[0].shape

    # ----------------------------------------------------------------------
    # HDF5 input and output (iohdf5)
    def _io(self, kind):
        return [io for io in self.block.io if isinstance(io, iohdf5) and io.iotype == kind]

    def _read_inputs(self, problem):
        """Arrays of the iotype "read" objects, with their halos: {name: (array, info)}
        (the datasets declared by ops_decl_dat_hdf5, opsc.py:702-705)."""
        from . import hdf5
        nd = problem["ndim"]
        out = {}
        for io in self._io("read"):
            path = io.read_file_name()
            for nm in _array_names(io.arrays, nd):
                a, info = hdf5.decl_dat_hdf5(path, self.block.blockname, hdf5.dataset_name(nm, self.block.blocknumber))
                if info["ndim"] != nd or list(info["size"]) != list(problem["n"]):
                    raise ValueError("%s in %s has size %s; the block has %s"
                                     % (nm, path, info["size"], list(problem["n"])))
                out[nm] = (a, info)
        return out

    def _defined_arrays(self, nd):
        """Arrays of the block an iohdf5 write may name: the conservative
        arrays, the DataObjects of the initial condition and the arrays read."""
        names = {"rho", "rhoE"} | {"rhou%d" % d for d in range(nd)}
        names |= {k for k in self._fields[1] if _KINDS.get(k) == "data"}
        if self.scheme.get("curvilinear"):
            names |= set(_metric_names(nd))
        names |= {"D%d%d" % (d, d) for d in range(nd) if self.scheme["stretch"][d]}
        if self._stores_teno():
            names.add("TENO")
        return names | set(self._inputs)

    def _stores_teno(self):
        """LLFTeno(formulation='adaptive', store_sensor=True) keeps its cut-off
        in the TENO dataset (teno.py:439-442)."""
        tenos = [s for s in self.block.discretisation_schemes.values() if isinstance(s, LLFTeno)]
        return bool(self.scheme.get("sensor") and tenos and tenos[0].store_sensor)

    def _init_equations(self):
        katz = [e for e in self.block.equation_classes if isinstance(e, Initialise_Katzer)]
        if katz:
            return list(katz[0].coordinate_evaluations)
        return list(flatten([e.equations for e in self.block.equation_classes
                             if isinstance(e, GridBasedInitialisation)]))

    def _write(self, io, fname, blk, problem):
        """One iohdf5 write (ops_fetch_block_hdf5_file + ops_fetch_dat_hdf5_file,
        io_hdf5.py:99-127): conservative arrays from the device with the halo
        layers the boundary conditions fill, arrays of the initial condition
        evaluated over the same widened grid (as the initialisation kernel
        covers the halos), arrays read from HDF5 as read."""
        from . import hdf5
        nd = problem["ndim"]
        cons = ["rho"] + ["rhou%d" % d for d in range(nd)] + ["rhoE"]
        bn, bname = self.block.blocknumber, self.block.blockname
        hdf5.fetch_block_hdf5_file(fname, bname, bn, nd)
        widened = None
        metric = _metric_names(nd) if self.scheme.get("curvilinear") else []
        for nm in _array_names(io.arrays, nd):
            if nm in cons:
                a, dm, dp = blk.get_field_halo(cons.index(nm))
            elif nm in metric:
                # the metric arrays hold the periodic images in their halos
                # (metric_boundary_condition, metric.py:315-341)
                dm, dp = blk.halo_depths()
                a = np.pad(blk.metric_field(metric.index(nm)),
                           [(-dm[d], dp[d]) for d in reversed(range(nd))], mode="wrap")
            elif re.fullmatch(r"D(\d)\1", nm) and int(nm[1]) < nd and problem["stretch"][int(nm[1])]:
                # D_dd of a stretched direction: mirrored across its faces by the
                # metric boundary kernel (D[-t] = D[t], metric.py:359-379), a
                # function of its own index elsewhere
                s_ = int(nm[1])
                dm, dp = blk.halo_depths()
                a = blk.metric_field(s_ * nd + s_)
                for d in range(nd):
                    ax = nd - 1 - d
                    width = [(0, 0)] * nd
                    width[ax] = (-dm[d], dp[d])
                    a = np.pad(a, width, mode="reflect" if d == s_ else "edge")
            elif nm == "TENO":
                # the adaptive cut-off of the last stage at every interior point
                # (teno.py:430-443); the halo layers are left at zero
                dm, dp = blk.halo_depths()
                a = np.pad(_adaptive_ct(blk.sensor(), problem["teno_a1"], problem["teno_a2"]),
                           [(-dm[d], dp[d]) for d in reversed(range(nd))])
            elif nm in self._fields[1]:
                dm, dp = blk.halo_depths()
                if widened is None:
                    env, shape = self._env(problem, dm, dp)
                    widened = {}
                    self._evaluate(self._init_equations(), env, shape, widened)
                a = widened[nm]
            else:
                a, info = self._inputs[nm]
                dm, dp = info["d_m"], info["d_p"]
            hdf5.fetch_dat_hdf5_file(fname, bname, bn, hdf5.dataset_name(nm, bn), a, dm, dp)

    @staticmethod
    def _evaluate(equations, env, shape, out, forbid=()):
        for eq in equations:
            rhs = eq.rhs
            syms = sorted(rhs.free_symbols, key=lambda s: s.name)
            bad = [s for s in syms if s.name in forbid]
            if bad:
                raise NotImplementedError("boundary values depending on the solution (%s) are not implemented "
                                          "by the HIPC target" % bad)
            vals = [_lookup(env, s) for s in syms]
            missing = [s for s, v in zip(syms, vals) if v is None]
            if missing:
                raise ValueError("equation uses undefined symbols %s" % missing)
            f = sympy.lambdify(syms, rhs, modules="numpy", dummify=True)
            val = np.broadcast_to(np.asarray(f(*vals), dtype=np.float64), shape).copy()
            env[eq.lhs] = val
            if not isinstance(eq.lhs, _GridSymbol):
                out[str(eq.lhs)] = val

    def initial_state(self, problem):
        """Evaluate the initial condition over the interior grid: the
        GridBasedInitialisation equations, or the Initialise_Katzer profile."""
        nd = problem["ndim"]
        shape = tuple(reversed(problem["n"]))
        names = ["rho"] + ["rhou%d" % d for d in range(nd)] + ["rhoE"]
        self._inputs = self._read_inputs(problem)
        env, _ = self._env(problem)
        out = {}
        katz = [e for e in self.block.equation_classes if isinstance(e, Initialise_Katzer)]
        inits = [e for e in self.block.equation_classes if isinstance(e, GridBasedInitialisation)]
        if katz:
            k = katz[0]
            self._evaluate(k.coordinate_evaluations, env, shape, out)
            wd = k.wall_direction
            if wd != 1:
                raise NotImplementedError("the boundary-layer initialisation expects the wall normal in x1")
            if "x%d" % wd not in out:
                raise ValueError("Initialise_Katzer needs the coordinate DataObject(x%d)" % wd)
            from .boundary_layer import SimilarityProfile
            bl = SimilarityProfile(k.xMach, k.Re, Pr=0.72, gama=1.4, Tw=-1.0, Tinf=k.Tinf)
            rho, u, rv, T = bl.evaluate(out["x%d" % wd])
            gama, Minf = problem["gama"], problem.get("Minf", 1.0)
            out["rho"], out["rhou0"], out["rhou1"] = rho, rho * u, rv
            if nd == 3:
                out["rhou2"] = np.zeros(shape)
            out["rhoE"] = rho * T / (gama * (gama - 1) * Minf ** 2) + 0.5 * (out["rhou0"] ** 2 + rv ** 2) / rho
        elif inits:
            self._evaluate(list(flatten([e.equations for e in inits])), env, shape, out)
        elif not all(nm in self._inputs for nm in names):
            raise ValueError("no GridBasedInitialisation on the block")
        for nm in names:
            if nm not in out and nm in self._inputs:
                # a restart: the conservative array read from HDF5
                out[nm] = np.ascontiguousarray(env[sympy.Symbol(nm)])
            if nm not in out:
                raise ValueError("initial condition does not define %s" % nm)
        self._fields = (env, out)
        return np.stack([out[nm] for nm in names])

    def _attach_boundary_data(self, problem, q0):
        """Coordinates of stretched directions and the planes of Dirichlet and
        inlet faces, from the initial-condition evaluation."""
        nd = problem["ndim"]
        env, fields = self._fields
        shape = tuple(reversed(problem["n"]))
        coords = {}
        for d in range(nd):
            if problem["stretch"][d]:
                name = "x%d" % d
                if name not in fields:
                    raise ValueError("a stretched direction needs its coordinate DataObject(%s) in the "
                                     "initialisation" % name)
                line = np.moveaxis(fields[name], nd - 1 - d, -1).reshape(-1, problem["n"][d])[0]
                coords[d] = np.ascontiguousarray(line)
        if coords:
            problem["coords"] = coords
        if problem.get("curvilinear"):
            # the coordinates over the interior and the 2 points beyond it the
            # metric's central derivatives read (the initial condition covers
            # the halos, gridbasedinit.py:46-57)
            h = 2
            env2, shape2 = self._env(problem, [-h] * nd, [h] * nd)
            wide = {}
            self._evaluate(self._init_equations(), env2, shape2, wide)
            missing = ["x%d" % d for d in range(nd) if "x%d" % d not in wide]
            if missing:
                raise ValueError("a curvilinear block needs its coordinate DataObjects %s in the initialisation"
                                 % missing)
            problem["grid_coords"] = np.ascontiguousarray(np.stack([wide["x%d" % d] for d in range(nd)]))
            problem["grid_halo"] = h
        planes = {}
        conserved = {"rho", "rhoE"} | {"rhou%d" % d for d in range(nd)}
        for d in range(nd):
            for s in range(2):
                bc = self.block.boundaries[d][s]
                if bc.kind not in ("dirichlet", "inlet_pextrap", "isothermal_wall"):
                    continue
                if bc.kind == "isothermal_wall":
                    eqs = [e for e in bc.equations if str(e.lhs) == "rhoE"]
                    want = sympy.Symbol("rho") * sympy.Symbol("Twall") / (
                        sympy.Symbol("gama") * (sympy.Symbol("gama") - 1.0) * sympy.Symbol("Minf") ** 2.0)
                    if len(eqs) != 1 or sympy.simplify(eqs[0].rhs - want) != 0:
                        raise NotImplementedError("the isothermal wall implements rhoE = rho*Twall/"
                                                  "(gama*(gama-1)*Minf**2) only")
                    continue
                if bc.kind == "inlet_pextrap":
                    src = q0
                else:
                    out = {}
                    env2 = dict(env)
                    self._evaluate(bc.equations, env2, shape, out, forbid=conserved)
                    missing = [nm for nm in ["rho"] + ["rhou%d" % k for k in range(nd)] + ["rhoE"] if nm not in out]
                    if missing:
                        raise ValueError("DirichletBC on direction %d side %d does not set %s" % (d, s, missing))
                    src = np.stack([out[nm] for nm in ["rho"] + ["rhou%d" % k for k in range(nd)] + ["rhoE"]])
                face = np.take(src, 0 if s == 0 else -1, axis=1 + (nd - 1 - d))
                e1, e2 = [e for e in range(3) if e != d][:2]
                n1 = problem["n"][e1] if e1 < nd else 1
                n2 = problem["n"][e2] if e2 < nd else 1
                planes[(d, s)] = np.ascontiguousarray(face.reshape(nd + 2, n2, n1))
        if planes:
            problem["planes"] = planes
        binom = [f for f in self.block.equation_classes if isinstance(f, BinomialFilter)]
        if binom and binom[0].grid_condition is not None:
            # the grid condition over the interior, from the coordinates and grid indices
            cond = sympy.sympify(binom[0].grid_condition)
            syms = sorted(cond.free_symbols, key=lambda z: z.name)
            look = dict(env)
            look.update({sympy.Symbol(k): val for k, val in fields.items()})
            vals = [_lookup(look, z) for z in syms]
            missing = [z for z, v in zip(syms, vals) if v is None]
            if missing:
                raise ValueError("grid_condition uses undefined symbols %s" % missing)
            f = sympy.lambdify(syms, cond, modules="numpy", dummify=True)
            mask = np.broadcast_to(np.asarray(f(*vals), dtype=bool), shape)
            problem["filters"]["binomial"]["mask"] = np.ascontiguousarray(mask, dtype=np.uint8)
        return problem

    def run(self, niter=None, return_block=False):
        """Evaluate the initial condition (reading the iohdf5 inputs), run
        niter iterations and write the iohdf5 outputs: after every iteration k
        with k % save_every == 0 and after the loop (algorithm.py:443-465)."""
        problem = self.problem()
        q0 = self.initial_state(problem)
        problem = self._attach_boundary_data(problem, q0)
        nd = problem["ndim"]
        writes = self._io("write")
        defined = self._defined_arrays(nd)
        for io in writes:
            io.check_datasets(defined | {"rhou"})
            io.file_name()
            if io.save_every:
                io.file_name(int(io.save_every))
        n = int(self.values.get("niter", 0) if niter is None else niter)
        names = _state_names(nd)
        mon = getattr(self.alg, "simulation_monitor", None)
        every, nan_name = getattr(self, "print_every", None), getattr(self, "nan_check", None)
        if nan_name is not None and str(nan_name) not in names:
            raise NotImplementedError("print_iteration_ops checks a conservative array %s, not %s" % (names, nan_name))
        # states after k steps that something reads: the writes after step k;
        # the monitor after step 1 and every frequency-th step; the iteration
        # print of step k at the top of its loop body, i.e. after k-1 steps
        saves = {k for io in writes if io.save_every for k in range(int(io.save_every), n + 1, int(io.save_every))}
        monitored = ({k for k in range(mon.frequency, n + 1, mon.frequency)} | ({1} if n >= 1 else set())
                     if mon is not None else set())
        printed = {k - 1 for k in range(every, n + 1, every)} if every else set()
        out = open(mon.output_file, "w") if mon is not None and mon.output_file else None
        blk = native.Block(problem)
        dt = float(problem["dt"])

        def nan_check(name, k):
            if blk.nan_count(names.index(str(name))):
                raise FloatingPointError("NaN detected in %s after iteration %d" % (name, k))

        try:
            blk.set_state(q0)
            done = 0
            for k in sorted(saves | monitored | printed):
                blk.run(k - done)
                done = k
                if k in printed:
                    print("Iteration is %d" % (k + 1))
                    if nan_name is not None:
                        nan_check(nan_name, k)
                if k in monitored:
                    if mon.NaNcheck:
                        nan_check(mon.arrays[0], k)
                    text = ([mon.header()] if k == 1 else []) + [
                        mon.line(k, k * dt, [blk.point(names.index(a), loc)
                                             for a, loc in zip(mon.arrays, mon.probe_locations)])]
                    for t in text:
                        if out is not None:
                            out.write(t + "\n")
                        else:
                            print(t)
                for io in writes:
                    if io.save_every and k in saves and k % int(io.save_every) == 0:
                        self._write(io, io.file_name(k), blk, problem)
            blk.run(n - done)
            for io in writes:
                self._write(io, io.file_name(), blk, problem)
            q = blk.get_state()
        except BaseException:
            blk.close()
            raise
        finally:
            if out is not None:
                out.close()
        result = {nm: q[i] for i, nm in enumerate(names)}
        if return_block:
            return result, blk
        blk.close()
        return result


def print_iteration_ops(simulation_name="opensbli", every=250, NaN_check=None):
    """Print 'Iteration is k' at the top of every iteration k that is a
    multiple of `every`, with a NaN check of the conservative array named by
    NaN_check there (helperfunctions.py:172-190 inserts the same into the
    generated time loop; ops_NaNcheck stops the run, here FloatingPointError)."""
    prog = _CURRENT.get(simulation_name)
    if prog is None:
        raise ValueError("no HIPC program named %r: call HIPC(alg) first" % simulation_name)
    prog.print_every = int(every)
    prog.nan_check = None if NaN_check is None else str(NaN_check)
    return prog


def substitute_simulation_parameters(constants, values, simulation_name="opensbli"):
    """Give values to the simulation constants (helperfunctions.py:130-149).
    Values are C expressions as in the reference (e.g. '2*M_PI/block0np0')."""
    prog = _CURRENT.get(simulation_name)
    if prog is None:
        raise ValueError("no HIPC program named %r: call HIPC(alg) first" % simulation_name)
    prog.set_values(constants, values)
    return prog


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("math", "re", "np", "sympy", "native")]
