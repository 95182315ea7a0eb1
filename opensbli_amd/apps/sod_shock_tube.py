"""1-D Sod shock tube through the OpenSBLI-style API, as the reference's app
(apps/Sod_shock_tube/Sod_shock_tube.py): the Euler equations with LLF TENO5
(Roe average), RungeKuttaLS(3), Dirichlet faces holding the left
(rho, u, p) = (1, 0, 1) and right (0.125, 0, 0.1) states, the diaphragm at
x = 0.5, x = i/(n - 1); the reference runs 200 points, dt = 2e-4 to t = 0.2
and plots the density against the exact solution it ships (reference.txt).
"""
from opensbli_amd.api import (HIPC, ConstituentRelations, DataObject, DirichletBC, Double, EinsteinEquation,
                              GridBasedInitialisation, GridVariable, LLFTeno, RoeAverage, RungeKuttaLS,
                              SimulationBlock, SimulationDataType, SimulationEquations, TraditionalAlgorithmRK,
                              iohdf5, parse_expr, substitute_simulation_parameters)

NDIM = 1


def build(n=200, dt="0.0002", niter="ceil(0.2/0.0002)", output=True):
    sc1 = "**{'scheme':'Teno'}"
    eqs = ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc1,
           "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)  )" % sc1,
           "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s)  )" % sc1]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)"]
    ee = EinsteinEquation()
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in eqs:
        sim.add_equations(ee.expand(t, NDIM, "x", [], ["gama"]))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", [], ["gama"]))
    block = SimulationBlock(NDIM, block_number=0)
    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    cons = ["Eq(DataObject(rho), d)", "Eq(DataObject(rhou0), d*u0)", "Eq(DataObject(rhoE), p/(gama-1.0) + 0.5* d *(u0**2.0))"]
    ic = ["Eq(DataObject(x0), block.deltas[0]*block.grid_indexes[0])",
          "Eq(GridVariable(d), Piecewise((1.0, DataObject(x0) < 0.5), (0.125, True)))",
          "Eq(GridVariable(p), Piecewise((1.0, DataObject(x0) < 0.5), (0.10, True)))",
          "Eq(GridVariable(u0), Piecewise((0.0, DataObject(x0) < 0.5),(0, True)))"] + cons
    init = GridBasedInitialisation()
    init.add_equations([parse_expr(e, local_dict=env) for e in ic])
    side_state = lambda d, p: [parse_expr(e, local_dict=env) for e in  # noqa: E731
                               ["Eq(GridVariable(d), %r)" % d, "Eq(GridVariable(u0), 0.0)",
                                "Eq(GridVariable(p), %r)" % p] + cons]
    block.set_block_boundaries([DirichletBC(0, 0, side_state(1.0, 1.0)), DirichletBC(0, 1, side_state(0.125, 0.1))])
    llf = LLFTeno(5, averaging=RoeAverage([0, 1]))
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
    substitute_simulation_parameters(["gama", "dt", "niter", "block0np0", "Delta0block0", "eps", "TENO_CT"],
                                     ["1.4", dt, str(niter), str(n), "1.0/(block0np0-1)", "1.0e-16", "1.0e-5"])
    return prog
