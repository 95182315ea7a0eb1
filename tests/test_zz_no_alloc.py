"""The allocation-free adjoint step (osbl_adjoint_step_dev_noalloc) against
the allocating one (osbl_adjoint_step_dev) through the C-ABI: the same sums
in the same order on the same inputs (a 2-D WENO5-Z step runs its x and y
VJPs in one launch in the allocating entry point and in two here, with the y
part added last either way), so the results agree bitwise.
Two calls in a row, the second on a smaller block (the fixed scratch then
holds the first call's contents), then the other schemes, a 3-D viscous
characteristic block, and a block too large for the fixed scratch, which the
entry point refuses before launching anything."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(problem, q0, seed):
    """One RK3 step's stage inputs from q0, then one adjoint step of a seeded
    lambda through each entry point; returns both results (numpy)."""
    import torch
    from opensbli_amd.native import Block
    b = Block(problem)
    b.set_state(q0)
    ns = b.num_stages()
    stages = torch.empty((ns,) + q0.shape, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    b.step_stages_dev(stages.data_ptr())
    lam = torch.tensor(np.random.RandomState(seed).normal(size=q0.shape), dtype=torch.float64, device="cuda")
    la, lb = lam.clone(), lam.clone()
    torch.cuda.synchronize()
    b.adjoint_step_dev(stages.data_ptr(), la.data_ptr())
    b.adjoint_step_dev_noalloc(stages.data_ptr(), lb.data_ptr())
    b.sync()
    out = la.cpu().numpy(), lb.cpu().numpy()
    b.close()
    return out


def _kh(n, conv, viscous=False):
    from opensbli_amd.problems import kelvin_helmholtz
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    problem, q0 = kelvin_helmholtz(n, amps, conv=conv)
    if viscous:
        problem.update(viscous="const", Re=200.0, Pr=0.72, Minf=0.5)
    return problem, q0


def _check(label, a, b):
    d = np.abs(a - b).max()
    _say(f"{label}: max |noalloc - alloc| = {d:.3e}, max |alloc| = {np.abs(a).max():.3e}")
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b), (label, d)


def test_noalloc_adjoint_step_two_calls_second_smaller():
    _say("case noalloc adjoint step: KH 64^2 WENO-Z viscous, then KH 32^2 WENO-Z viscous")
    p, q = _kh(64, "weno_z", viscous=True)
    _check("KH 64^2 weno_z viscous", *_both(p, q, 1))
    p, q = _kh(32, "weno_z", viscous=True)
    _check("KH 32^2 weno_z viscous", *_both(p, q, 2))


@pytest.mark.parametrize("conv,viscous", [("central_skew", False), ("central_skew", True), ("weno_z", False)])
def test_noalloc_adjoint_step_schemes(conv, viscous):
    _say(f"case noalloc adjoint step: KH 48^2 {conv} viscous={viscous}")
    p, q = _kh(48, conv, viscous)
    _check(f"KH 48^2 {conv} viscous={viscous}", *_both(p, q, 3))


@pytest.mark.parametrize("conv", ["weno_z", "central_skew"])
def test_noalloc_adjoint_step_3d_viscous(conv):
    from opensbli_amd.problems import taylor_green
    _say(f"case noalloc adjoint step: TGV 16^3 {conv} viscous")
    p, q = taylor_green(16, conv=conv)
    rs = np.random.RandomState(7)
    q = q * (1.0 + 0.01 * rs.uniform(-1, 1, q.shape))
    _check(f"TGV 16^3 {conv} viscous", *_both(p, q, 4))


def test_noalloc_adjoint_step_refuses_oversized_block():
    import torch
    from opensbli_amd.native import Block, OsblError
    from opensbli_amd.problems import taylor_green
    _say("case noalloc adjoint step: TGV 112^3 WENO-Z needs more than the fixed scratch")
    p, q = taylor_green(112, conv="weno_z")
    b = Block(p)
    b.set_state(q)
    lam = torch.zeros((5,) + q.shape[1:], dtype=torch.float64, device="cuda")
    stages = torch.zeros((3, 5) + q.shape[1:], dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(OsblError, match="fixed scratch"):
        b.adjoint_step_dev_noalloc(stages.data_ptr(), lam.data_ptr())
    b.sync()
    assert float(lam.abs().max()) == 0.0
    b.close()
