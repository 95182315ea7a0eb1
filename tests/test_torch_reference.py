"""The PyTorch restatement used as the autograd reference of the adjoint
agrees with the oracle (CPU, float64)."""
import numpy as np
import torch

from conftest import rel_l2


def test_torch_reference_matches_oracle(oracle):
    import torch_reference as tr
    from opensbli_amd.problems import kelvin_helmholtz
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    problem, q0 = kelvin_helmholtz(32, amps, conv="central_skew")
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    r.run(20)
    qt = tr.integrate(torch.tensor(q0, dtype=torch.float64), 20, problem["dt"], problem["delta"], problem["gama"])
    assert np.all(rel_l2(qt.numpy(), r.get_state()) < 1e-12)


def test_torch_weno_z_matches_oracle(oracle):
    """The WENO5-Z restatement: RHS and 20 RK3 steps of the KH layer."""
    import torch_reference as tr
    from opensbli_amd.problems import kelvin_helmholtz
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    problem, q0 = kelvin_helmholtz(32, amps, conv="weno_z")
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    rr = r.rhs()
    rt = tr.rhs_weno_z(torch.tensor(q0, dtype=torch.float64), problem["delta"], problem["gama"]).numpy()
    assert np.all(rel_l2(rt, rr) < 1e-11), rel_l2(rt, rr)
    r.set_state(q0)
    r.run(20)
    qt = tr.integrate_weno(torch.tensor(q0, dtype=torch.float64), 20, problem["dt"], problem["delta"],
                           problem["gama"])
    assert np.all(rel_l2(qt.numpy(), r.get_state()) < 1e-11)


def test_torch_viscous_matches_oracle(oracle):
    """The viscous restatement (constant viscosity) on the KH layer (2-D,
    central skew and WENO5-Z) and on the TGV (3-D): RHS, and 10 RK3 steps of
    the 2-D central case."""
    import torch_reference as tr
    from opensbli_amd.problems import kelvin_helmholtz, taylor_green
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    for conv in ("central_skew", "weno_z"):
        problem, q0 = kelvin_helmholtz(32, amps, conv=conv)
        problem.update(viscous="const", Re=200.0, Pr=0.72, Minf=0.5)
        r = oracle.RefSolver(problem)
        r.set_state(q0)
        rr = r.rhs()
        rt = tr.rhs_ns(torch.tensor(q0, dtype=torch.float64), problem["delta"], problem["gama"], 0.5, 200.0, 0.72,
                       conv).numpy()
        assert np.all(rel_l2(rt, rr) < 1e-12), (conv, rel_l2(rt, rr))
    problem, q0 = kelvin_helmholtz(32, amps, conv="central_skew")
    problem.update(viscous="const", Re=200.0, Pr=0.72, Minf=0.5)
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    r.run(10)
    qt = tr.integrate_ns(torch.tensor(q0, dtype=torch.float64), 10, problem["dt"], problem["delta"], problem["gama"],
                         0.5, 200.0, 0.72)
    assert np.all(rel_l2(qt.numpy(), r.get_state()) < 1e-12)
    problem, q0 = taylor_green(12, conv="central_skew")
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    rr = r.rhs()
    rt = tr.rhs_ns(torch.tensor(q0, dtype=torch.float64), problem["delta"], problem["gama"], problem["Minf"],
                   problem["Re"], problem["Pr"]).numpy()
    assert np.all(rel_l2(rt, rr) < 1e-12), rel_l2(rt, rr)
