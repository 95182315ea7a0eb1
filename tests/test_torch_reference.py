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
