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
    GridBasedInitialisation            gridbasedinit.py:15
    TraditionalAlgorithmRK             algorithm.py:336
    substitute_simulation_parameters   helperfunctions.py:130

Instead of expanding the Einstein notation symbolically, HIPC recognises the
equation systems the GPU kernels implement (the compressible Euler/Navier-
Stokes equations in skew-symmetric or conservative form, the constituent
relations and the constant-viscosity stress tensor and heat flux of the
reference's apps) and refuses anything else with the reference's error types.
The initial condition equations are evaluated on the host with SymPy.
"""
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


def GridVariable(name):
    """Kernel-local variable of an initial condition (grid.py:214)."""
    return _mk(name, "grid")


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

    @property
    def time_advance_arrays(self):
        return ["rho", "rhou", "rhoE"]


class ConstituentRelations(_EquationSet):
    """Constituent relations (opensbliequations.py:444)."""


class GridBasedInitialisation(_EquationSet):
    """Initial-condition equations evaluated over the grid before the time
    loop (gridbasedinit.py:15-67); here evaluated on the host with SymPy."""


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
        if formulation is not None:
            raise NotImplementedError("adaptive TENO is not supported by the HIPC target yet")
        self.conv = "teno5" if order == 5 else "teno6"
        self.averaging = (averaging or SimpleAverage([0, 1])).kind
        self.name = "TenoDerivative"
        self.schemetype = "Spatial"


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


class ExtrapolationBC:
    """Extrapolation boundary (extrapolation.py:8); zeroth order."""

    def __init__(self, direction, side, order, scheme=None, plane=True):
        if order != 0:
            raise NotImplementedError("the HIPC target implements zeroth-order extrapolation only")
        self.direction, self.side, self.kind = direction, side, "extrap0"


# ---------------------------------------------------------------------------
# I/O and documentation hooks of the reference, accepted and recorded
# ---------------------------------------------------------------------------
class iohdf5:
    """HDF5 output placement (io_hdf5.py:19).  The HIPC target returns the
    final arrays from HIPC.run(); it records the request and writes nothing."""

    def __init__(self, save_every=None, **kwargs):
        self.arrays = []
        self.save_every = save_every
        self.kwargs = kwargs

    def add_arrays(self, arrays):
        self.arrays.extend(flatten(list(arrays)) if isinstance(arrays, (list, tuple)) else [arrays])


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


VISCOUS = {"rhou_i": {(1, "Der(tau_i_j,x_j)")}, "rhoE": {(1, "Der(u_i*tau_i_j,x_j)")}}
STRESS = "Eq(tau_i_j,(1.0/Re)*(Der(u_i,x_j)+Der(u_j,x_i)-(2/3)*KD(_i,_j)*Der(u_k,x_k)))"
HEAT_POS = "Eq(q_j,(1.0/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
HEAT_NEG = "Eq(q_j,(-1.0/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
CONSTITUENT = {
    "u_i": "rhou_i/rho",
    "p": "(gama-1)*(rhoE-rho*(1/2)*(KD(_i,_j)*u_i*u_j))",
    "a": "(gama*p/rho)**0.5",
    "T": "p*gama*Minf*Minf/(rho)",
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
        terms[m.group(1)] = set(_split_terms(rhs))
        subs.update(_norm(s) for s in e.substitutions)
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
    viscous_terms = VISCOUS
    has_visc = any(t in terms[k] for k, v in viscous_terms.items() for t in v)
    heat = None
    if has_visc:
        for sign in (1, -1):
            if (sign, "Der(q_j,x_j)") in terms["rhoE"]:
                heat = sign
    conv = None
    averaging = "roe"
    for name, want in [("skew", SKEW), ("weno", _cons("@Weno")), ("teno", _cons("@Teno")), ("central", _cons(""))]:
        ok = True
        for var in ("rho", "rhou_i", "rhoE"):
            got = set(terms[var])
            if has_visc:
                got -= viscous_terms.get(var, set())
                if var == "rhoE" and heat is not None:
                    got.discard((heat, "Der(q_j,x_j)"))
            if got != want[var]:
                ok = False
        if ok:
            conv = name
            break
    if conv is None:
        raise NotImplementedError("equation system not recognised by the HIPC target: %s" % terms)
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
        for var, want in VISCOUS.items():
            if not want <= terms[var]:
                raise NotImplementedError("partial viscous terms are not supported")
        if _norm(STRESS) not in subs:
            raise NotImplementedError("only the constant-viscosity stress tensor (1.0/Re)*(...) is supported")
        heat_ok = (heat == 1 and _norm(HEAT_POS) in subs) or (heat == -1 and _norm(HEAT_NEG) in subs)
        if not heat_ok:
            raise NotImplementedError("heat flux form not supported")
        viscous = "const"
    # constituent relations
    have = {}
    for c in crs:
        for e in c.equations:
            lhs, rhs = _parse_eq(e.text)
            have[lhs] = rhs
    need = ["u_i", "p"] + (["a"] if conv not in ("central_skew", "central_cons") else []) + \
        (["T"] if viscous != "none" else [])
    for k in need:
        if k not in have:
            raise ValueError("Constituent relation is not found for %s" % k)
        if have[k] != _norm(CONSTITUENT[k]):
            raise NotImplementedError("constituent relation for %s not supported: %s" % (k, have[k]))
    return dict(conv=conv, averaging=averaging, viscous=viscous, rk=temporal[0].rk)


# ---------------------------------------------------------------------------
# the target
# ---------------------------------------------------------------------------
_CURRENT = {}


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
                 bc=[[b.boundaries[d][s].kind for s in range(2)] for d in range(nd)], **self.scheme)
        for k in ("Minf", "Re", "Pr", "eps"):
            if k in v:
                p[k] = float(v[k])
        if "TENO_CT" in v:
            p["teno_ct"] = float(v["TENO_CT"])
        return p

    def initial_state(self, problem):
        """Evaluate the GridBasedInitialisation equations over the interior grid."""
        inits = [e for e in self.block.equation_classes if isinstance(e, GridBasedInitialisation)]
        if not inits:
            raise ValueError("no GridBasedInitialisation on the block")
        nd = problem["ndim"]
        shape = tuple(reversed(problem["n"]))
        env = {}
        for k, val in self.values.items():
            env[sympy.Symbol(k)] = val
        grids = np.meshgrid(*[np.arange(n) for n in reversed(problem["n"])], indexing="ij")
        for d in range(nd):
            env[sympy.Symbol("idx%d" % d)] = grids[nd - 1 - d].astype(np.float64)
        env[sympy.pi] = math.pi
        out = {}
        for eq in flatten([e.equations for e in inits]):
            rhs = eq.rhs
            syms = sorted(rhs.free_symbols, key=lambda s: s.name)
            missing = [s for s in syms if s not in env]
            if missing:
                raise ValueError("initial condition uses undefined symbols %s" % missing)
            f = sympy.lambdify(syms, rhs, modules="numpy")
            val = np.broadcast_to(np.asarray(f(*[env[s] for s in syms]), dtype=np.float64), shape).copy()
            env[eq.lhs] = val
            out[str(eq.lhs)] = val
        names = ["rho"] + ["rhou%d" % d for d in range(nd)] + ["rhoE"]
        for nm in names:
            if nm not in out:
                raise ValueError("initial condition does not define %s" % nm)
        return np.stack([out[nm] for nm in names])

    def run(self, niter=None, return_block=False):
        problem = self.problem()
        q0 = self.initial_state(problem)
        blk = native.Block(problem)
        blk.set_state(q0)
        n = int(self.values.get("niter", 0) if niter is None else niter)
        blk.run(n)
        q = blk.get_state()
        names = ["rho"] + ["rhou%d" % d for d in range(problem["ndim"])] + ["rhoE"]
        result = {nm: q[i] for i, nm in enumerate(names)}
        if return_block:
            return result, blk
        blk.close()
        return result


def substitute_simulation_parameters(constants, values, simulation_name="opensbli"):
    """Give values to the simulation constants (helperfunctions.py:130-149).
    Values are C expressions as in the reference (e.g. '2*M_PI/block0np0')."""
    prog = _CURRENT.get(simulation_name)
    if prog is None:
        raise ValueError("no HIPC program named %r: call HIPC(alg) first" % simulation_name)
    prog.set_values(constants, values)
    return prog


__all__ = [n for n in dir() if not n.startswith("_") and n not in ("math", "re", "np", "sympy", "native")]
