"""Differentiable time integration: the hand-written discrete adjoint of the
GPU kernels behind a torch.autograd.Function, with checkpoint-every-k
recomputation.

    qN = integrate(q0, problem, nsteps, every=20)      # q0: float64 CUDA tensor (nvar, ...)
    J = kinetic_energy(qN, problem); J.backward()       # q0.grad = dJ/dq0

Forward: the block is advanced with the GPU kernels; the state is kept every
`every` steps, and the input states of every RK stage of the last segment
are stored as it runs.  Backward: for each segment, from last to first, the
segment is recomputed from its checkpoint (except the last) while the input
states of every RK stage are stored, then the adjoint of each step (osbl_adjoint_step_dev) is applied
from the last step back; the recomputation of the next segment runs on a
second context of the same problem, on its own stream, during the adjoint
steps of the current one.  The block is then left at the forward output.
Covered: the central skew-symmetric scheme and the
LLF characteristic WENO5-Z scheme, with or without constant-viscosity terms,
on periodic blocks with RungeKutta(3) (TENO's cut-off is not differentiable).
"""
import math

import torch

from . import native


def _sync():
    torch.cuda.synchronize()


class _Integrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, block, nsteps, every):
        q0c = q0.detach().contiguous()
        _sync()
        block.set_state_dev([q0c[m].data_ptr() for m in range(block.nvar)])
        checkpoints = [q0c.clone()]
        # the last segment's stage inputs are stored as it runs (the backward
        # pass starts there and would otherwise recompute them first, with
        # nothing to overlap); the buffer has room for the longest segment and
        # becomes one of the backward pass's two stage buffers
        last = None
        done = 0
        while done < nsteps:
            k = min(every, nsteps - done)
            if done + k == nsteps and ctx.needs_input_grad[0]:
                last = torch.empty((min(every, nsteps), block.num_stages()) + tuple(q0c.shape),
                                   dtype=torch.float64, device=q0c.device)
                for i in range(k):
                    block.step_stages_dev(last[i].data_ptr())
            else:
                block.run(k)
            done += k
            snap = torch.empty_like(q0c)
            block.get_state_dev([snap[m].data_ptr() for m in range(block.nvar)])
            block.sync()
            checkpoints.append(snap)
        ctx.block = block
        ctx.nsteps = nsteps
        ctx.every = every
        ctx.checkpoints = checkpoints
        ctx.last = last
        return checkpoints[-1].clone()

    @staticmethod
    def backward(ctx, gN):
        block = ctx.block
        nv = block.nvar
        lam = gN.detach().to(torch.float64).contiguous().clone()
        nseg = len(ctx.checkpoints) - 1
        ns = block.num_stages()

        def steps_of(seg):
            return min(ctx.every, ctx.nsteps - seg * ctx.every)

        kmax = max(steps_of(seg) for seg in range(nseg))
        last, ctx.last = ctx.last, None
        bufs = [last if last is not None else
                torch.empty((kmax, ns) + tuple(lam.shape), dtype=torch.float64, device=lam.device),
                torch.empty((kmax, ns) + tuple(lam.shape), dtype=torch.float64, device=lam.device)]
        # a second context of the same problem recomputes the stage inputs of
        # the next segment (backwards) on its own stream while this block runs
        # the adjoint steps of the current one; the two share only read-only
        # checkpoints and alternate between the two stage buffers
        twin = _twin(block)

        def restart(seg):
            q = ctx.checkpoints[seg]
            twin.set_state_dev([q[m].data_ptr() for m in range(nv)])

        _sync()
        if last is None:
            restart(nseg - 1)
            for s in range(steps_of(nseg - 1)):
                twin.step_stages_dev(bufs[0][s].data_ptr())
            twin.sync()
        cur = 0
        for seg in range(nseg - 1, -1, -1):
            # the two streams' launches are issued alternately, so that
            # neither waits for the host to finish queueing the other's
            nrec = steps_of(seg - 1) if seg > 0 else 0
            if nrec:
                restart(seg - 1)
            k = steps_of(seg)
            for i in range(max(k, nrec)):
                if i < nrec:
                    twin.step_stages_dev(bufs[1 - cur][i].data_ptr())
                if i < k:
                    block.adjoint_step_dev(bufs[cur][k - 1 - i].data_ptr(), lam.data_ptr())
            block.sync()
            twin.sync()
            cur = 1 - cur
        # leave the block at the forward output, where the forward pass left
        # it, so a caller can go on integrating the same block
        qN = ctx.checkpoints[-1]
        block.set_state_dev([qN[m].data_ptr() for m in range(nv)])
        block.sync()
        return lam, None, None, None


def _twin(block):
    """The second context used to recompute segments during the backward
    pass (made once per block, kept with it)."""
    twin = getattr(block, "_adjoint_twin", None)
    if twin is None:
        twin = native.Block(block.p)
        block._adjoint_twin = twin
    return twin


def integrate(q0, block, nsteps, every=20):
    """Advance q0 (float64 CUDA tensor, (nvar, [nz,] [ny,] nx)) by nsteps RK
    iterations of `block` (a native.Block); differentiable with respect to q0."""
    if q0.dtype != torch.float64 or not q0.is_cuda:
        raise TypeError("q0 must be a float64 CUDA tensor")
    return _Integrate.apply(q0, block, int(nsteps), int(every))


def kinetic_energy(q, problem):
    """J = sum 1/2 |rho u|^2 / rho * dV over the block."""
    nd = problem["ndim"]
    dv = 1.0
    for d in range(nd):
        dv *= problem["delta"][d]
    rho = q[0]
    mom2 = sum(q[1 + d] ** 2 for d in range(nd))
    return 0.5 * (mom2 / rho).sum() * dv


def kh_initial_state(amplitudes, n, sigma=0.05, gama=1.4):
    """Kelvin-Helmholtz initial state as a differentiable function of the 16
    mode amplitudes (torch, float64, on the amplitudes' device); the same
    formula as problems.kelvin_helmholtz_state."""
    dev = amplitudes.device
    d = 1.0 / n
    x = torch.arange(n, dtype=torch.float64, device=dev) * d
    y, xx = torch.meshgrid(x, x, indexing="ij")
    inner = (y >= 0.25) & (y < 0.75)
    rho = torch.where(inner, torch.full_like(y, 2.0), torch.full_like(y, 1.0))
    u = torch.where(inner, torch.full_like(y, 0.5), torch.full_like(y, -0.5))
    env = torch.exp(-(y - 0.25) ** 2 / (2 * sigma ** 2)) + torch.exp(-(y - 0.75) ** 2 / (2 * sigma ** 2))
    m = torch.arange(1, amplitudes.shape[0] + 1, dtype=torch.float64, device=dev)
    modes = torch.sin(2.0 * math.pi * m[:, None, None] * xx[None]) * env[None]
    v = (amplitudes[:, None, None] * modes).sum(0)
    p = 2.5
    return torch.stack([rho, rho * u, rho * v, p / (gama - 1.0) + 0.5 * rho * (u * u + v * v)])
