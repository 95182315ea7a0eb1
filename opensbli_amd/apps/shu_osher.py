"""1-D Shu-Osher shock/entropy-wave interaction through the OpenSBLI-style
API, as the reference's app (apps/shu_osher/shu_osher.py): the Euler
equations with LLF WENO5-Z (Roe average), RungeKuttaLS(3) on [-5, 5]; a
Mach 3 shock at x = -4 moving into rho = 1 + 0.2 sin(5x); Dirichlet faces
holding the initial condition evaluated at x = -5 and x = 5.  The reference
runs 3200 points, dt = 2e-4 to t = 1.8.
"""
import sympy

from opensbli_amd.api import (HIPC, ConstituentRelations, DataObject, DirichletBC, Double, EinsteinEquation,
                              GridBasedInitialisation, GridVariable, LLFWeno, RoeAverage, RungeKuttaLS,
                              SimulationBlock, SimulationDataType, SimulationEquations, TraditionalAlgorithmRK,
                              iohdf5, parse_expr, substitute_simulation_parameters)

NDIM = 1


def build(n=3200, dt="0.0002", niter="ceil(1.8/0.0002)", output=True):
    sc1 = "**{'scheme':'Weno'}"
    eqs = ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc1,
           "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)  )" % sc1,
           "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s)  )" % sc1]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)"]
    ee = EinsteinEquation()
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in eqs:
        sim.add_equations(ee.expand(t, NDIM, "x", [], ["gama", "Minf"]))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", [], ["gama", "Minf"]))
    block = SimulationBlock(NDIM, block_number=0)
    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    x0 = "Eq(GridVariable(x0), -5.0 + block.deltas[0]*block.grid_indexes[0])"
    rest = ["Eq(GridVariable(u0), Piecewise((2.629369, x0<-4.0),(0.0, x0>=-4.0),(0.0,True)))",
            "Eq(GridVariable(p), Piecewise((10.33333, x0<-4.0),(1.0, x0>=-4.0),(0.0,True)))",
            "Eq(GridVariable(d), Piecewise((3.857143, x0<-4.0),(1.0+0.2*sin(5*x0), x0>=-4.0),(0,True)))",
            "Eq(DataObject(rho), d)", "Eq(DataObject(rhou0), d*u0)",
            "Eq(DataObject(rhoE), p/(gama-1.0) + 0.5* d *(u0**2.0))"]
    init = GridBasedInitialisation()
    init.add_equations([parse_expr(e, local_dict=env) for e in [x0, "Eq(DataObject(x0), GridVariable(x0))"] + rest])
    bnd = [parse_expr(e, local_dict=env) for e in [x0] + rest]
    # the boundary states: the initial condition with x0 = -5 and x0 = 5 (shu_osher.py:83-87)
    left = [e.subs({sympy.Symbol("x0"): -5.0}) for e in bnd]
    right = [e.subs({sympy.Symbol("x0"): 5.0}) for e in bnd]
    block.set_block_boundaries([DirichletBC(0, 0, left), DirichletBC(0, 1, right)])
    llf = LLFWeno(5, formulation="Z", averaging=RoeAverage([0, 1]))
    rk = RungeKuttaLS(3)
    block.set_discretisation_schemes({llf.name: llf, rk.name: rk})
    if output:
        h5 = iohdf5(**{"iotype": "Write"})
        h5.add_arrays(sim.time_advance_arrays)
        h5.add_arrays([DataObject("x0")])
        block.setio([h5])
    block.set_equations([cr, sim, init])
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)
    substitute_simulation_parameters(["gama", "Minf", "dt", "niter", "block0np0", "Delta0block0"],
                                     ["1.4", "0.1", dt, str(niter), str(n), "10.0/(block0np0-1)"])
    return prog
