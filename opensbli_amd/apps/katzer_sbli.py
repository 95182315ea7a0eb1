"""2-D shock-wave/boundary-layer interaction of Katzer (Mach 2, Re = 950)
through the OpenSBLI-style API, as the reference's app
(apps/katzer_SBLI/katzer_SBLI.py): conservative Navier-Stokes with the
adaptive TENO5 (Roe average, modified Ducros sensor, the cut-off stored in the
TENO array), Sutherland viscosity, StoreSome(4) viscous terms, the sinh grid
stretched normal to the wall (MetricsEquation), InletPressureExtrapolateBC /
zeroth-order extrapolation in x0, an isothermal wall (Twall = 1.67619431) and
on top a Dirichlet shock generator (the free stream for x0 <= 40, the state
behind an oblique shock beyond), ReducedAccess closures, the Initialise_Katzer
similarity profile, RungeKuttaLS(3, formulation='SSP') (katzer_SBLI.py:72).
The reference runs 500 x 250 points,
dt = 0.04, 250000 iterations, and compares the skin friction and the wall
pressure with reference_data.txt (apps/katzer_SBLI/plot.py).
"""
from opensbli_amd.api import (HIPC, ConstantObject, ConstituentRelations, DataObject, DirichletBC, Double,
                              EinsteinEquation, ExtrapolationBC, GridVariable, InletPressureExtrapolateBC,
                              Initialise_Katzer, IsothermalWallBC, LLFTeno, MetricsEquation, ReducedAccess,
                              RoeAverage, RungeKuttaLS, ShockSensor, SimulationBlock, SimulationDataType,
                              SimulationEquations, StoreSome, TraditionalAlgorithmRK, iohdf5, parse_expr,
                              substitute_simulation_parameters)

NDIM = 2


def build(n=(500, 250), niter=250000, dt="0.04", output=True):
    """The Katzer SBLI program on an n[0] x n[1] grid of [0, 400] x [0, 115]."""
    sc1 = "**{'scheme':'Teno'}"
    mass = "Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc1
    momentum = "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s) + Der(tau_i_j,x_j) )" % sc1
    energy = "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s) - Der(q_j,x_j) + Der(u_i*tau_i_j ,x_j) )" % sc1
    stress_tensor = "Eq(tau_i_j, (mu/Re)*(Der(u_i,x_j)+ Der(u_j,x_i) - (2/3)* KD(_i,_j)* Der(u_k,x_k)))"
    heat_flux = "Eq(q_j, (-mu/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"
    substitutions = [stress_tensor, heat_flux]
    constants = ["Re", "Pr", "gama", "Minf", "SuthT", "RefT"]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)", "Eq(T, p*gama*Minf*Minf/(rho))",
            "Eq(mu, (T**(1.5)*(1.0+SuthT/RefT)/(T+SuthT/RefT)))"]
    ee = EinsteinEquation()
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in (mass, momentum, energy):
        sim.add_equations(ee.expand(t, NDIM, "x", substitutions, constants))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", substitutions, constants))

    block = SimulationBlock(NDIM, block_number=0)
    metriceq = MetricsEquation()
    metriceq.generate_transformations(NDIM, "x", [(False, False), (True, False)], 2)
    sim.apply_metrics(metriceq)
    shock_sensor, sensor_array = ShockSensor().ducros_equations(block, "x", metriceq)
    cr.add_equations(shock_sensor)
    llf = LLFTeno(5, formulation="adaptive", averaging=RoeAverage([0, 1]), sensor=sensor_array, store_sensor=True)
    cent = StoreSome(4, "u0 u1 u2 T")
    rk = RungeKuttaLS(3, formulation="SSP")
    block.set_discretisation_schemes({llf.name: llf, cent.name: cent, rk.name: rk})

    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    for con in ("gama", "Minf", "Twall", "Lx1", "by"):
        env[con] = ConstantObject(con)
    x_loc = parse_expr("Eq(GridVariable(x0), block.deltas[0]*block.grid_indexes[0])", local_dict=env)
    rhoE_wall = parse_expr("Eq(DataObject(rhoE), DataObject(rho)*Twall/(gama*(gama-1.0)*Minf**2.0))",
                           local_dict=env)
    top = [parse_expr(e, local_dict=env) for e in (
        "Eq(DataObject(rho), Piecewise((1.129734572, (x0)>40.0), (1.00000596004, True)))",
        "Eq(DataObject(rhou0), Piecewise((1.0921171, (x0)>40.0), (1.00000268202, True)))",
        "Eq(DataObject(rhou1), Piecewise((-0.058866065, (x0)>40.0), (0.00565001630205, True)))",
        "Eq(DataObject(rhoE), Piecewise((1.0590824, (x0)>40.0), (0.94644428042, True)))")]
    block.set_block_boundaries([
        InletPressureExtrapolateBC(0, 0, scheme=ReducedAccess()),
        ExtrapolationBC(0, 1, order=0, scheme=ReducedAccess()),
        IsothermalWallBC(1, 0, [rhoE_wall], scheme=ReducedAccess()),
        DirichletBC(1, 1, [x_loc] + top, scheme=ReducedAccess()),
    ])
    gridx0 = parse_expr("Eq(DataObject(x0), block.deltas[0]*block.grid_indexes[0])", local_dict=env)
    gridx1 = parse_expr("Eq(DataObject(x1), Lx1*sinh(by*block.deltas[1]*block.grid_indexes[1]/Lx1)/sinh(by))",
                        local_dict=env)
    initial = Initialise_Katzer([(False, DataObject("x0")), (True, DataObject("x1"))], 50, 950.0, 2.0, 288.0,
                                [gridx0, gridx1])
    if output:
        h5 = iohdf5(**{"iotype": "Write"})
        h5.add_arrays(sim.time_advance_arrays)
        h5.add_arrays([DataObject("x0"), DataObject("x1"), DataObject("D11"), DataObject("TENO")])
        block.setio(h5)
    block.set_equations([cr, sim, initial, metriceq])
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)
    substitute_simulation_parameters(
        ["gama", "Minf", "Pr", "Re", "Twall", "dt", "niter", "block0np0", "block0np1", "Delta0block0",
         "Delta1block0", "SuthT", "RefT", "eps", "TENO_CT", "Lx1", "by", "teno_a1", "teno_a2", "epsilon"],
        ["1.4", "2.0", "0.72", "950.0", "1.67619431", dt, str(niter), str(n[0]), str(n[1]), "400.0/(block0np0-1)",
         "115.0/(block0np1-1)", "110.4", "288.0", "1e-15", "1e-5", "115.0", "5.0", "10.5", "4.5", "1.0e-30"])
    return prog


if __name__ == "__main__":
    import sys
    prog = build(niter=int(sys.argv[1]) if len(sys.argv) > 1 else 1000)
    out = prog.run()
    print({k: float(abs(v).max()) for k, v in out.items()})
