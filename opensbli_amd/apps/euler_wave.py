"""2-D Euler density wave through the OpenSBLI-style API, as the reference's
app (apps/euler_wave/euler_wave.py): LLF WENO5-Z with the simple average,
RungeKutta(3), periodic on [0, 2)^2; rho = 1 + 0.2 sin(pi (x0 + x1)),
u = (1, -0.5), p = 1, whose exact solution is the translation
rho(x - t u) (apps/euler_wave/plot.py:93-103).  The reference runs 400^2
points, dt = 0.001, 2500 iterations.
"""
from opensbli_amd.api import (HIPC, ConstituentRelations, DataObject, Double, EinsteinEquation,
                              GridBasedInitialisation, GridVariable, LLFWeno, PeriodicBC, RungeKutta, SimpleAverage,
                              SimulationBlock, SimulationDataType, SimulationEquations, TraditionalAlgorithmRK,
                              iohdf5, parse_expr, substitute_simulation_parameters)

NDIM = 2


def build(n=400, dt="0.001", niter=2500, output=True):
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
        sim.add_equations(ee.expand(t, NDIM, "x", [], ["gama"]))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", [], ["gama"]))
    block = SimulationBlock(NDIM, block_number=0)
    llf = LLFWeno(5, formulation="Z", averaging=SimpleAverage([0, 1]))
    rk = RungeKutta(3)
    block.set_discretisation_schemes({llf.name: llf, rk.name: rk})
    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    ic = ["Eq(DataObject(x0), block.deltas[0]*block.grid_indexes[0])",
          "Eq(DataObject(x1), block.deltas[1]*block.grid_indexes[1])",
          "Eq(GridVariable(d), 1+0.2*sin(pi*(DataObject(x0)+DataObject(x1))))",
          "Eq(GridVariable(u0), 1.0)", "Eq(GridVariable(u1), -0.5)", "Eq(GridVariable(p), 1.0)",
          "Eq(DataObject(rho), d)", "Eq(DataObject(rhou0), d*u0)", "Eq(DataObject(rhou1), d*u1)",
          "Eq(DataObject(rhoE), p/(gama-1) + 0.5* d *(u0**2+u1**2))"]
    init = GridBasedInitialisation()
    init.add_equations([parse_expr(e, local_dict=env) for e in ic])
    block.set_block_boundaries([PeriodicBC(d, s) for d in range(NDIM) for s in (0, 1)])
    if output:
        h5 = iohdf5(**{"iotype": "Write"})
        h5.add_arrays(sim.time_advance_arrays)
        h5.add_arrays([DataObject("x0"), DataObject("x1")])
        block.setio([h5])
    block.set_equations([cr, sim, init])
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)
    substitute_simulation_parameters(["gama", "dt", "niter", "block0np0", "block0np1", "Delta0block0", "Delta1block0"],
                                     ["1.4", dt, str(niter), str(n), str(n), "2.0/(block0np0)", "2.0/(block0np1)"])
    return prog
