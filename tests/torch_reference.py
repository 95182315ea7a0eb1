"""Plain PyTorch FP64 statement of the inviscid central skew-symmetric scheme
with RungeKutta(3) on a periodic 2-D grid (test infrastructure): the autograd
reference for the hand-written adjoint.

The same discretisation as the reference's TGV-style Skew form
(parsing.py:75-111, scheme.py:187-252) without viscous terms: 4th-order
central first derivatives with periodic wrap, 3-register RK3
(rk_sbli.py:58-60,117-133).
"""
import torch


def d1(f, axis, inv):
    """4th-order central first derivative along a tensor axis (periodic)."""
    return inv * ((1.0 / 12.0) * (torch.roll(f, 2, axis) - torch.roll(f, -2, axis))
                  + (2.0 / 3.0) * (torch.roll(f, -1, axis) - torch.roll(f, 1, axis)))


def rhs(q, delta, gama):
    nd = q.dim() - 1
    nv = nd + 2
    rho = q[0]
    u = [q[1 + d] / rho for d in range(nd)]
    p = (gama - 1.0) * (q[nv - 1] - 0.5 * rho * sum(ui * ui for ui in u))
    r = [torch.zeros_like(rho) for _ in range(nv)]
    for j in range(nd):
        axis = nd - 1 - j  # tensor axis of coordinate direction j (x fastest)
        inv = 1.0 / delta[j]
        uj = u[j]
        duj = d1(uj, axis, inv)
        for m in range(nv):
            r[m] = r[m] - 0.5 * (d1(q[m] * uj, axis, inv) + uj * d1(q[m], axis, inv) + q[m] * duj)
        r[1 + j] = r[1 + j] - d1(p, axis, inv)
        r[nv - 1] = r[nv - 1] - d1(p * uj, axis, inv)
    return torch.stack(r)


def rk3_step(q, dt, delta, gama):
    rknew = [2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0]
    rkold = [1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0]
    qold = q
    cur = q
    for s in range(3):
        r = rhs(cur, delta, gama)
        cur = qold + dt * rknew[s] * r
        qold = qold + dt * rkold[s] * r
    return cur


def integrate(q, nsteps, dt, delta, gama):
    for _ in range(nsteps):
        q = rk3_step(q, dt, delta, gama)
    return q
