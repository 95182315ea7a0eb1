"""2-D inviscid oblique shock reflection through the OpenSBLI-style API, as
the reference's app (apps/inviscid_shock_reflection/inviscid_shock.py): the
Euler equations with LLF WENO5-Z (simple average) and RungeKuttaLS(3) on
[0, 350] x [0, 115]; a Mach 2 inflow (Dirichlet) on the left, zeroth-order
extrapolation on the right, a symmetry plane at the bottom, and on top a
Dirichlet face holding the inflow state for x < 40 and the state behind an
incident shock (p2/p1 = 1.18646663) for x >= 40.  The shock reflects
regularly off the bottom; the reference runs 457 x 255 points, dt = 0.1,
10000 iterations.
"""
from opensbli_amd.api import (HIPC, ConstantObject, ConstituentRelations, DataObject, DirichletBC, Double,
                              EinsteinEquation, ExtrapolationBC, GridBasedInitialisation, GridVariable, LLFWeno,
                              RungeKuttaLS, SimpleAverage, SimulationBlock, SimulationDataType, SimulationEquations,
                              SymmetryBC, TraditionalAlgorithmRK, iohdf5, parse_expr,
                              substitute_simulation_parameters)

NDIM = 2
# the state behind the incident shock (inviscid_shock.py:108-121)
POST_SHOCK = dict(rho=1.129734572, u0=0.9667023804242755, u1=-0.052106102140246795, p_ratio=1.18646663)


def build(n=(457, 255), niter=10000, dt="0.1", output=True):
    """The shock reflection program on an n[0] x n[1] grid."""
    sc1 = "**{'scheme':'Weno'}"
    mass = "Eq(Der(rho,t), - %s)" % ("Conservative(rhou_j,x_j,%s)" % sc1)
    momentum = "Eq(Der(rhou_i,t) , -%s  )" % ("Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)" % sc1)
    energy = "Eq(Der(rhoE,t), - %s  )" % ("Conservative((p+rhoE)*u_j,x_j, %s)" % sc1)
    consts = ["gama", "Minf"]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)"]
    ee = EinsteinEquation()
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in (mass, momentum, energy):
        sim.add_equations(ee.expand(t, NDIM, "x", [], consts))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", [], consts))

    llf = LLFWeno(5, formulation="Z", averaging=SimpleAverage([0, 1]))
    rk = RungeKuttaLS(3)
    block = SimulationBlock(NDIM, block_number=0)
    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    for c in consts:
        env[c] = ConstantObject(c)
    x0 = parse_expr("Eq(DataObject(x0), block.deltas[0]*block.grid_indexes[0])", local_dict=env)
    x1 = parse_expr("Eq(DataObject(x1), block.deltas[1]*block.grid_indexes[1])", local_dict=env)
    d_in = parse_expr("Eq(GridVariable(d), 1.0)", local_dict=env)
    u0_in = parse_expr("Eq(GridVariable(u0), 1.0)", local_dict=env)
    u1_in = parse_expr("Eq(GridVariable(u1), 0.0)", local_dict=env)
    p_in = parse_expr("Eq(GridVariable(p), 1.0/(gama*Minf**2.0))", local_dict=env)
    rho = parse_expr("Eq(DataObject(rho), d)", local_dict=env)
    rhou0 = parse_expr("Eq(DataObject(rhou0), d*u0)", local_dict=env)
    rhou1 = parse_expr("Eq(DataObject(rhou1), d*u1)", local_dict=env)
    rhoE = parse_expr("Eq(DataObject(rhoE), p/(gama-1) + 0.5* d *(u0**2+u1**2))", local_dict=env)
    init = GridBasedInitialisation()
    init.add_equations([x0, x1, d_in, u0_in, u1_in, p_in, rho, rhou0, rhou1, rhoE])

    pre = 1.0 / (1.4 * 4.0)
    post = pre * POST_SHOCK["p_ratio"]
    shock_loc = parse_expr("Eq(GridVariable(shock_loc), 40.0)", local_dict=env)
    xg = parse_expr("Eq(GridVariable(x), DataObject(x0))", local_dict=env)
    d = parse_expr("Eq(GridVariable(d), Piecewise((1.0 ,x<shock_loc),(%r, x>=shock_loc),(1.0,True)))"
                   % POST_SHOCK["rho"], local_dict=env)
    u0 = parse_expr("Eq(GridVariable(u0), Piecewise((1.0 ,x<shock_loc),(%r, x>=shock_loc),(1.0,True)))"
                    % POST_SHOCK["u0"], local_dict=env)
    u1 = parse_expr("Eq(GridVariable(u1), Piecewise((0.0 ,x<shock_loc),(%r, x>=shock_loc),(1.0,True)))"
                    % POST_SHOCK["u1"], local_dict=env)
    p = parse_expr("Eq(GridVariable(p), Piecewise( (%f,x<shock_loc),(%f, x>=shock_loc),(1.0,True)))" % (pre, post),
                   local_dict=env)
    block.set_block_boundaries([
        DirichletBC(0, 0, [d_in, u0_in, u1_in, p_in, rho, rhou0, rhou1, rhoE]),
        ExtrapolationBC(0, 1, order=0),
        SymmetryBC(1, 0),
        DirichletBC(1, 1, [xg, shock_loc, d, u0, u1, p, rho, rhou0, rhou1, rhoE]),
    ])
    if output:
        h5 = iohdf5(**{"iotype": "Write"})
        h5.add_arrays(sim.time_advance_arrays)
        h5.add_arrays([DataObject("x0"), DataObject("x1")])
        block.setio([h5])
    block.set_equations([cr, sim, init])
    block.set_discretisation_schemes({llf.name: llf, rk.name: rk})
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)
    substitute_simulation_parameters(
        ["gama", "Minf", "dt", "niter", "block0np0", "block0np1", "Delta0block0", "Delta1block0"],
        ["1.4", "2.0", dt, str(niter), str(n[0]), str(n[1]), "350.0/(block0np0-1)", "115.0/(block0np1-1)"])
    return prog


if __name__ == "__main__":
    import sys
    prog = build(niter=int(sys.argv[1]) if len(sys.argv) > 1 else 1000)
    out = prog.run()
    print({k: float(abs(v).sum()) for k, v in out.items()})
