"""The discrete adjoint of the constant-viscosity terms (the reverse of the
stress tensor and heat flux, taylor_green_vortex.py:14-19), with the central
skew-symmetric and the WENO5-Z convective residuals, on periodic blocks:
vector-Jacobian products against PyTorch FP64 autograd of the restatement in
tests/torch_reference.py (itself pinned to the oracle), a dot-product test
against finite differences of the GPU residual, and the gradient of the
final kinetic energy over RK3 steps against autograd."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VISC = dict(viscous="const", Re=200.0, Pr=0.72, Minf=0.5)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _generic_state(n, seed=0):
    rs = np.random.RandomState(seed)
    y, x = np.meshgrid(np.arange(n) / n, np.arange(n) / n, indexing="ij")
    k = rs.uniform(1, 3, 8)
    rho = 1.5 + 0.3 * np.sin(2 * np.pi * (x + k[0] * y)) + 0.01 * rs.rand(n, n)
    u = 0.3 * np.cos(2 * np.pi * (k[1] * x - y)) + 0.01 * rs.rand(n, n)
    v = 0.2 * np.sin(2 * np.pi * (x + 2 * y)) + 0.01 * rs.rand(n, n)
    p = 2.5 + 0.2 * np.cos(2 * np.pi * (2 * x + y)) + 0.01 * rs.rand(n, n)
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


def _kh_visc(n, conv):
    from opensbli_amd.problems import kelvin_helmholtz
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    problem, q0 = kelvin_helmholtz(n, amps, conv=conv)
    problem.update(VISC)
    return problem, q0, amps


def _gpu_vjp(b, q0, mu):
    import torch
    qd = torch.tensor(q0, dtype=torch.float64, device="cuda").contiguous()
    md = torch.tensor(mu, dtype=torch.float64, device="cuda").contiguous()
    out = torch.zeros_like(qd)
    torch.cuda.synchronize()
    b.rhs_vjp_dev(qd.data_ptr(), md.data_ptr(), out.data_ptr())
    b.sync()
    return out.cpu().numpy()


@pytest.mark.parametrize("conv", ["central_skew", "weno_z"])
def test_viscous_vjp_vs_torch_autograd_2d(conv):
    import torch
    import torch_reference as tr
    from opensbli_amd.native import Block
    _say(f"case viscous adjoint vjp vs autograd 32^2 {conv}")
    problem, _, _ = _kh_visc(32, conv)
    q0 = _generic_state(32)
    mu = np.random.RandomState(2).normal(size=q0.shape)
    b = Block(problem)
    gk = _gpu_vjp(b, q0, mu)
    qt = torch.tensor(q0, dtype=torch.float64, requires_grad=True)
    r = tr.rhs_ns(qt, problem["delta"], problem["gama"], VISC["Minf"], VISC["Re"], VISC["Pr"], conv)
    (gt,) = torch.autograd.grad((r * torch.tensor(mu)).sum(), qt)
    err = np.abs(gk - gt.numpy()).max() / np.abs(gt.numpy()).max()
    _say("max rel diff", err)
    assert err < 1e-10


def test_viscous_vjp_vs_torch_autograd_3d():
    import torch
    import torch_reference as tr
    from opensbli_amd.native import Block
    from opensbli_amd.problems import taylor_green
    _say("case viscous adjoint vjp vs autograd TGV 16^3 central skew")
    problem, q0 = taylor_green(16, conv="central_skew")
    rs = np.random.RandomState(3)
    q0 = q0 * (1.0 + 0.01 * rs.uniform(-1, 1, q0.shape))
    q0[3] += 0.01 * rs.uniform(-1, 1, q0[3].shape)
    mu = rs.normal(size=q0.shape)
    b = Block(problem)
    gk = _gpu_vjp(b, q0, mu)
    qt = torch.tensor(q0, dtype=torch.float64, requires_grad=True)
    r = tr.rhs_ns(qt, problem["delta"], problem["gama"], problem["Minf"], problem["Re"], problem["Pr"])
    (gt,) = torch.autograd.grad((r * torch.tensor(mu)).sum(), qt)
    err = np.abs(gk - gt.numpy()).max() / np.abs(gt.numpy()).max()
    _say("max rel diff", err)
    assert err < 1e-10


def test_viscous_vjp_dot_product():
    from opensbli_amd.native import Block
    _say("case viscous adjoint vjp dot product 32^2 central skew")
    problem, _, _ = _kh_visc(32, "central_skew")
    q0 = _generic_state(32, seed=5)
    b = Block(problem)
    rs = np.random.RandomState(6)
    dq = rs.normal(size=q0.shape) * np.abs(q0).mean(axis=(1, 2))[:, None, None] * 1e-3
    mu = rs.normal(size=q0.shape)
    eps = 1e-4

    def R(q):
        b.set_state(q)
        return b.rhs()

    fd = (R(q0 + eps * dq) - R(q0 - eps * dq)) / (2 * eps)
    lhs = float(np.sum(mu * fd))
    rhs_ = float(np.sum(_gpu_vjp(b, q0, mu) * dq))
    _say("dot products", lhs, rhs_)
    assert abs(lhs - rhs_) < 1e-7 * abs(lhs)


@pytest.mark.parametrize("conv", ["central_skew", "weno_z"])
def test_viscous_gradient_vs_torch_autograd(conv):
    """Gradient of the final kinetic energy of the viscous KH layer (64^2,
    20 RK3 steps, checkpoints every 10) with respect to the 16 mode
    amplitudes, against autograd of the restatement."""
    import torch
    import torch_reference as tr
    from opensbli_amd.adjoint import integrate, kh_initial_state, kinetic_energy
    from opensbli_amd.native import Block
    _say(f"case viscous adjoint gradient vs autograd 64^2 20 steps {conv}")
    n, steps = 64, 20
    problem, _, amps = _kh_visc(n, conv)
    b = Block(problem)
    a = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
    J = kinetic_energy(integrate(kh_initial_state(a, n), b, steps, 10), problem)
    J.backward()
    g = a.grad.cpu().numpy()
    a2 = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
    qN = tr.integrate_ns(kh_initial_state(a2, n), steps, problem["dt"], problem["delta"], problem["gama"],
                         VISC["Minf"], VISC["Re"], VISC["Pr"], conv)
    Jt = kinetic_energy(qN, problem)
    Jt.backward()
    gt = a2.grad.cpu().numpy()
    _say("J", float(J), float(Jt), "max rel grad diff", np.abs(g - gt).max() / np.abs(gt).max())
    assert abs(float(J) - float(Jt)) < 1e-12 * abs(float(Jt))
    assert np.abs(g - gt).max() < 1e-8 * np.abs(gt).max()


def test_sutherland_adjoint_refused():
    from opensbli_amd.native import Block, OsblError
    problem, q0, _ = _kh_visc(32, "central_skew")
    problem.update(viscous="sutherland", SuthT=110.4, RefT=288.0)
    b = Block(problem)
    with pytest.raises(OsblError, match="Sutherland"):
        _gpu_vjp(b, q0, np.ones_like(q0))
