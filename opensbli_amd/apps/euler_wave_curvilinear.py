"""2-D Euler density wave on a wavy curvilinear mesh through the OpenSBLI-style
API, as the reference's app (apps/euler_wave_curvilinear/euler_wave.py): a
MetricsEquation curvilinear in both directions, the Euler equations in strong
conservation form in xi,
    Der(rho, t) = -Conservative(detJ rho U_j, xi_j)/detJ, ...,  U_i = D_i_j u_j,
LLF WENO5-Z with the Roe average, RungeKuttaLS(3), periodic in both
directions, the wavy mesh x0 = i dx + A sin(2 pi w) sin(6 pi j dy/2),
x1 = j dy + A sin(2 pi w) sin(6 pi i dx/2), and the wave
rho = 1 + 0.2 sin(pi (x0 + x1)), u0 = 1, u1 = -0.5, p = 1.  The conservative
arrays, the coordinates and the metric D00..D11 are written every save_every
iterations and at the end.
"""
from opensbli_amd.api import (HIPC, ConstituentRelations, DataObject, Double, EinsteinEquation,
                              GridBasedInitialisation, GridVariable, LLFWeno, MetricsEquation, PeriodicBC, RoeAverage,
                              RungeKuttaLS, SimulationBlock, SimulationDataType, SimulationEquations, StoreSome,
                              TraditionalAlgorithmRK, iohdf5, parse_expr, substitute_simulation_parameters)

NDIM = 2


def build(n=128, niter=5000, dt="0.0005", save_every=1000, amplitude=0.04):
    """The curvilinear Euler wave program on an n x n grid of [0, 2)^2."""
    metriceq = MetricsEquation()
    metriceq.generate_transformations(NDIM, "x", [(True, True), (True, True)], 2)
    sc = "**{'scheme':'Weno'}"
    mass = "Eq(Der(rho,t), - %s/detJ)" % ("Conservative(detJ * rho*U_j,xi_j,%s)" % sc)
    momentum = "Eq(Der(rhou_i,t) , -  %s/detJ)" % ("Conservative(detJ * (rhou_i*U_j + p*D_j_i), xi_j , %s)" % sc)
    energy = "Eq(Der(rhoE,t), - %s/detJ)" % ("Conservative(detJ * (p+rhoE)*U_j,xi_j, %s)" % sc)
    consts = ["gama", "Minf"]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE-rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)", "Eq(U_i, D_i_j*u_j)"]
    ee = EinsteinEquation()
    ee.optional_subs_dict = metriceq.metric_subs
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in (mass, momentum, energy):
        sim.add_equations(ee.expand(t, NDIM, "xi", [], consts))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "xi", [], consts))

    block = SimulationBlock(NDIM, block_number=0)
    llf = LLFWeno(5, formulation="Z", averaging=RoeAverage([0, 1]))
    cent = StoreSome(4, "u0 u1 T")
    rk = RungeKuttaLS(3)
    block.set_discretisation_schemes({llf.name: llf, cent.name: cent, rk.name: rk})

    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    wave = "%r*sin(2.0*pi*0.25)" % float(amplitude)
    ic = [
        "Eq(DataObject(x0), block.grid_indexes[0]*block.deltas[0] + %s*sin(6.0*pi*block.grid_indexes[1]*"
        "block.deltas[1]/2.0))" % wave,
        "Eq(DataObject(x1), block.grid_indexes[1]*block.deltas[1] + %s*sin(6.0*pi*block.grid_indexes[0]*"
        "block.deltas[0]/2.0))" % wave,
        "Eq(GridVariable(d), 1+0.2*sin(pi*(DataObject(x0)+DataObject(x1))))",
        "Eq(GridVariable(u0), 1.0)",
        "Eq(GridVariable(u1), -0.5)",
        "Eq(GridVariable(p), 1.0)",
        "Eq(DataObject(rho), d)",
        "Eq(DataObject(rhou0), d*u0)",
        "Eq(DataObject(rhou1), d*u1)",
        "Eq(DataObject(rhoE), p/(gama-1) + 0.5* d *(u0**2+u1**2))",
    ]
    init = GridBasedInitialisation()
    init.add_equations([parse_expr(e, local_dict=env) for e in ic])

    block.set_block_boundaries([PeriodicBC(d, s) for d in range(NDIM) for s in (0, 1)])
    h5 = iohdf5(save_every=save_every, **{"iotype": "Write"})
    h5.add_arrays(sim.time_advance_arrays)
    h5.add_arrays([DataObject("x0"), DataObject("x1"), DataObject("D00"), DataObject("D01"), DataObject("D10"),
                   DataObject("D11")])
    block.setio([h5])
    block.set_equations([cr, init, sim, metriceq])
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)
    substitute_simulation_parameters(
        ["gama", "dt", "niter", "block0np0", "block0np1", "Delta0block0", "Delta1block0"],
        ["1.4", dt, str(niter), str(n), str(n), "2.0/(block0np0)", "2.0/(block0np1)"])
    return prog


if __name__ == "__main__":
    import sys
    prog = build(int(sys.argv[1]) if len(sys.argv) > 1 else 128, niter=int(sys.argv[2]) if len(sys.argv) > 2 else 100,
                 save_every=50)
    out = prog.run()
    print({k: float(abs(v).sum()) for k, v in out.items()})
