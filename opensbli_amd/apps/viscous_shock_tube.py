"""2-D viscous shock tube (Daru & Tenaud; Re = 200) through the OpenSBLI-style
API, as the reference's app (apps/viscous_shock_tube/viscous_shock_tube.py):
the Navier-Stokes equations with LLF TENO5 (Roe average) and StoreSome(4)
viscous terms, constant viscosity mu/Re with mu a declared constant,
RungeKuttaLS(3); adiabatic walls (ReducedAccess closures) on the left, right
and bottom faces and a symmetry plane on top; the diaphragm at x = 0.5 between
rho = 120, p = 120/gama and rho = 1.2, p = 1.2/gama, at rest.  The reference
runs 500 x 250 points to t = 1 (dt = 5e-6, 200000 iterations) and compares the
bottom-wall density with reference_Re200.txt (Zhou et al., Phys. Fluids 30,
016102, 2018).
"""
from opensbli_amd.api import (HIPC, AdiabaticWallBC, ConstituentRelations, DataObject, Double, EinsteinEquation,
                              GridBasedInitialisation, GridVariable, LLFTeno, ReducedAccess, RoeAverage,
                              RungeKuttaLS, SimulationBlock, SimulationDataType, SimulationEquations, StoreSome,
                              SymmetryBC, TraditionalAlgorithmRK, iohdf5, parse_expr,
                              substitute_simulation_parameters)

NDIM = 2


def build(n=(500, 250), niter=200000, dt="0.000005", save_every=None, output=True):
    """The viscous shock tube program on an n[0] x n[1] grid of [0, 1] x [0, 0.5]."""
    sc1 = "**{'scheme':'Teno'}"
    mass = "Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc1
    momentum = "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s) + Der(tau_i_j,x_j) )" % sc1
    energy = "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s) - Der(q_j,x_j) + Der(u_i*tau_i_j ,x_j) )" % sc1
    stress_tensor = "Eq(tau_i_j, (mu/Re)*(Der(u_i,x_j)+ Der(u_j,x_i) - (2/3)* KD(_i,_j)* Der(u_k,x_k)))"
    heat_flux = "Eq(q_j, (-mu/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
    substitutions = [stress_tensor, heat_flux]
    constants = ["Re", "Pr", "gama", "Minf", "mu"]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)", "Eq(T, p*gama*Minf*Minf/(rho))"]
    ee = EinsteinEquation()
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in (mass, momentum, energy):
        sim.add_equations(ee.expand(t, NDIM, "x", substitutions, constants))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", substitutions, constants))

    block = SimulationBlock(NDIM, block_number=0)
    llf = LLFTeno(5, averaging=RoeAverage([0, 1]))
    cent = StoreSome(4, "u0 u1 u2")
    rk = RungeKuttaLS(3)
    block.set_discretisation_schemes({llf.name: llf, cent.name: cent, rk.name: rk})

    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    cond = [(120.0, 1.2), (0, 0), (0, 0), (120.0, 1.2)]
    ic = [
        "Eq(GridVariable(x_grid), block.deltas[0]*block.grid_indexes[0])",
        "Eq(DataObject(x0), GridVariable(x_grid))",
        "Eq(DataObject(x1), block.deltas[1]*block.grid_indexes[1])",
        "Eq(GridVariable(rho), Piecewise((%.15f, GridVariable(x_grid)<=0.5), (%.15f, True)))" % cond[0],
        "Eq(GridVariable(u0), Piecewise((%.15f, GridVariable(x_grid)<=0.5), (%.15f, True)))" % cond[1],
        "Eq(GridVariable(u1), Piecewise((%.15f, GridVariable(x_grid)<=0.5), (%.15f, True)))" % cond[2],
        "Eq(GridVariable(p), Piecewise((%.15f/gama, GridVariable(x_grid)<=0.5), (%.15f/gama, True)))" % cond[3],
        "Eq(DataObject(rho), GridVariable(rho))",
        "Eq(DataObject(rhou0), GridVariable(rho)*GridVariable(u0))",
        "Eq(DataObject(rhou1), GridVariable(rho)*GridVariable(u1))",
        "Eq(DataObject(rhoE), GridVariable(p)/(gama-1.0) + 0.5*GridVariable(rho)*(GridVariable(u0)**2 + "
        "GridVariable(u1)**2))",
    ]
    init = GridBasedInitialisation()
    init.add_equations([parse_expr(e, local_dict=env) for e in ic])

    block.set_block_boundaries([AdiabaticWallBC(0, 0, scheme=ReducedAccess()),
                                AdiabaticWallBC(0, 1, scheme=ReducedAccess()),
                                AdiabaticWallBC(1, 0, scheme=ReducedAccess()),
                                SymmetryBC(1, 1, scheme=ReducedAccess())])
    if output:
        h5 = iohdf5(save_every=save_every, **{"iotype": "Write"})
        h5.add_arrays(sim.time_advance_arrays)
        h5.add_arrays([DataObject("x0"), DataObject("x1")])
        block.setio([h5])
    block.set_equations([cr, sim, init])
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)
    substitute_simulation_parameters(
        ["mu", "gama", "Minf", "Pr", "Re", "dt", "niter", "block0np0", "block0np1", "Delta0block0", "Delta1block0",
         "eps", "TENO_CT"],
        ["1.0", "1.4", "1.0", "0.73", "200.0", dt, str(niter), str(n[0]), str(n[1]), "1.0/(block0np0-1)",
         "0.5/(block0np1-1)", "1e-15", "1e-6"])
    return prog


if __name__ == "__main__":
    import sys
    nx = int(sys.argv[1]) if len(sys.argv) > 1 else 500
    prog = build((nx, nx // 2), niter=int(sys.argv[2]) if len(sys.argv) > 2 else 1000)
    out = prog.run()
    print({k: float(abs(v).sum()) for k, v in out.items()})
