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
