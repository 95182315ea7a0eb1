"""2-D Kelvin-Helmholtz instability through the OpenSBLI-style API, as the
reference's app (apps/kelvin_helmholtz/kelvin_helmholtz.py): the Euler
equations in conservative form with LLF WENO5-Z (Roe average), RK3, periodic
in both directions; the shear-layer perturbation comes from the white-noise
array random_nums, written beforehand with output_hdf5 to data.h5 (with 5 halo
layers) and read back by an iohdf5 read.  The conservative arrays and the
coordinates are written every save_every iterations and at the end.
"""
import numpy as np

from opensbli_amd.api import (HIPC, ConstituentRelations, DataObject, Double, EinsteinEquation,
                              GridBasedInitialisation, GridVariable, LLFWeno, PeriodicBC, RoeAverage, RungeKutta,
                              SimulationBlock, SimulationDataType, SimulationEquations, TraditionalAlgorithmRK, iohdf5,
                              output_hdf5, parse_expr, substitute_simulation_parameters)

NDIM = 2


def build(n=512, niter=50000, dt="0.0001", save_every=10000, seed=None, noise_file="data.h5"):
    """The KH program on an n x n grid; writes the random_nums input file
    (numpy.random.RandomState(seed) when a seed is given) before returning."""
    sc = "**{'scheme':'Weno'}"
    eqs = ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc,
           "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)  )" % sc,
           "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s))" % sc]
    consts = ["gama", "Minf"]
    rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
            "Eq(a, (gama*p/rho)**0.5)"]
    ee = EinsteinEquation()
    sim = SimulationEquations()
    cr = ConstituentRelations()
    for t in eqs:
        sim.add_equations(ee.expand(t, NDIM, "x", [], consts))
    for t in rels:
        cr.add_equations(ee.expand(t, NDIM, "x", [], consts))

    block = SimulationBlock(NDIM, block_number=0)
    schemes = {}
    llf = LLFWeno(5, formulation="Z", averaging=RoeAverage([0, 1]))
    schemes[llf.name] = llf
    rk = RungeKutta(3)
    schemes[rk.name] = rk
    block.set_block_boundaries([PeriodicBC(d, s) for d in range(NDIM) for s in (0, 1)])
    block.set_discretisation_schemes(schemes)

    env = {"block": block, "GridVariable": GridVariable, "DataObject": DataObject}
    ic = [
        "Eq(DataObject(x0), -0.5+block.deltas[0]*block.grid_indexes[0])",
        "Eq(DataObject(x1), -0.5+block.deltas[1]*block.grid_indexes[1])",
        "Eq(GridVariable(x1c), DataObject(x1))",
        "Eq(GridVariable(d), Piecewise((1.0 ,x1c<-0.25),(1.0, x1c>0.25),(2.0,True)))",
        "Eq(GridVariable(u0),  Piecewise((-0.5 ,x1c<-0.25),(-0.5, x1c>0.25),(0.5+0.01*(DataObject(random_nums)-0.5),True)))",
        "Eq(GridVariable(u1),Piecewise((0 ,x1c<-0.25),(0, x1c>0.25),(0.01*(DataObject(random_nums)-0.5),True)))",
        "Eq(GridVariable(p), 2.5)",
        "Eq(DataObject(rho), d)",
        "Eq(DataObject(rhou0), d*u0)",
        "Eq(DataObject(rhou1), d*u1)",
        "Eq(DataObject(rhoE), p/(gama-1) + 0.5* d *(u0**2+u1**2))",
    ]
    init = GridBasedInitialisation()
    init.add_equations([parse_expr(e, local_dict=env) for e in ic])

    h5 = iohdf5(save_every=save_every, **{"iotype": "Write"})
    h5.add_arrays(sim.time_advance_arrays)
    h5.add_arrays([DataObject("x0"), DataObject("x1")])
    h5_read = iohdf5(**{"iotype": "Read", "filename": noise_file})
    h5_read.add_arrays([DataObject("random_nums")])
    block.setio([h5, h5_read])
    block.set_equations([cr, init, sim])
    block.discretise()
    alg = TraditionalAlgorithmRK(block)
    SimulationDataType.set_datatype(Double)
    prog = HIPC(alg)

    halos = [(-5, 5), (-5, 5)]
    rs = np.random if seed is None else np.random.RandomState(seed)
    noise = rs.rand(*[n + 10 for _ in range(NDIM)])
    output_hdf5(noise, "random_nums", halos, [n, n], block, filename=noise_file)

    substitute_simulation_parameters(
        ["gama", "Minf", "dt", "niter", "block0np0", "block0np1", "Delta0block0", "Delta1block0", "TENO_CT", "eps"],
        ["1.4", "2.0", dt, str(niter), str(n), str(n), "1.0/block0np0", "1.0/block0np1", "1e-5", "1e-15"])
    return prog


if __name__ == "__main__":
    import sys
    prog = build(int(sys.argv[1]) if len(sys.argv) > 1 else 512, niter=int(sys.argv[2]) if len(sys.argv) > 2 else 100,
                 save_every=50)
    out = prog.run()
    print({k: float(abs(v).sum()) for k, v in out.items()})
