"""Differentiable Acc and Phi on the MI355X kernels (torch.autograd).

``acc(pos, mass, eps2, mode)`` and ``phi(pos, mass, eps2, mode)`` evaluate
the self-gravitating pair sums of tupan/lib/src/acc_kernel_common.h:7-37 and
phi_kernel_common.h:7-31 (every particle acts on every particle, the r^2 > 0
mask drops self pairs) with positions as a (3, n) float64 device tensor (SoA
rows x, y, z) and masses / half-softenings as (n,) tensors.  Their backward
passes run the hand-written adjoint kernels (adjoint.hip): one further O(n^2)
pair sum that returns the gradients with respect to every position and mass
at once.  Softening is not differentiated.

``leapfrog_rollout`` integrates with the kick-drift-kick leapfrog (the
SIA21 kdk step of tupan/integrator/sia.py:322-337, one Acc evaluation per
step) and recomputes each segment of ``checkpoint_every`` steps in the
backward pass instead of storing it (torch.utils.checkpoint).
"""
import torch
from torch.utils.checkpoint import checkpoint

from . import device

_F5 = ("mass", "rx", "ry", "rz", "eps2")


def _fields(pos, mass, eps2):
    return {"mass": mass, "rx": pos[0], "ry": pos[1], "rz": pos[2],
            "eps2": eps2}


def _check(pos, mass, eps2):
    if pos.dim() != 2 or pos.shape[0] != 3:
        raise ValueError("pos must have shape (3, n)")
    n = pos.shape[1]
    for t, nm in ((mass, "mass"), (eps2, "eps2")):
        if t.shape != (n,):
            raise ValueError("%s must have shape (n,)" % nm)
    return pos.contiguous(), mass.contiguous(), eps2.contiguous()


class AccFunction(torch.autograd.Function):
    """a_k = -sum_j m_j r_kj / s_kj^3 for all k; backward via the Acc
    adjoint kernel."""

    @staticmethod
    def forward(ctx, pos, mass, eps2, mode):
        pos, mass, eps2 = _check(pos, mass, eps2)
        ps = _fields(pos, mass, eps2)
        out = torch.empty_like(pos)
        device.compute("acc", ps, ps, mode=mode, out=(out[0], out[1], out[2]))
        ctx.save_for_backward(pos, mass, eps2)
        ctx.mode = mode
        return out

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        pos, mass, eps2 = ctx.saved_tensors
        g = grad.contiguous()
        ps = _fields(pos, mass, eps2)
        ps.update({"gx": g[0], "gy": g[1], "gz": g[2]})
        gpos = torch.empty_like(pos)
        gm = torch.empty_like(mass)
        device.compute("acc_adjoint", ps, ps, mode=ctx.mode,
                       out=(gpos[0], gpos[1], gpos[2], gm))
        return gpos, gm, None, None


class PhiFunction(torch.autograd.Function):
    """phi_k = -sum_j m_j / s_kj; backward via the Phi adjoint kernel."""

    @staticmethod
    def forward(ctx, pos, mass, eps2, mode):
        pos, mass, eps2 = _check(pos, mass, eps2)
        ps = _fields(pos, mass, eps2)
        out = torch.empty_like(mass)
        device.compute("phi", ps, ps, mode=mode, out=(out,))
        ctx.save_for_backward(pos, mass, eps2)
        ctx.mode = mode
        return out

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        pos, mass, eps2 = ctx.saved_tensors
        ps = _fields(pos, mass, eps2)
        ps["h"] = grad.contiguous()
        gpos = torch.empty_like(pos)
        gm = torch.empty_like(mass)
        device.compute("phi_adjoint", ps, ps, mode=ctx.mode,
                       out=(gpos[0], gpos[1], gpos[2], gm))
        return gpos, gm, None, None


def acc(pos, mass, eps2, mode="fp64"):
    return AccFunction.apply(pos, mass, eps2, mode)


def phi(pos, mass, eps2, mode="fp64"):
    return PhiFunction.apply(pos, mass, eps2, mode)


def _segment(pos, vel, a, mass, eps2, dt, steps, mode):
    for _ in range(steps):
        vel = vel + a * (0.5 * dt)
        pos = pos + vel * dt
        a = acc(pos, mass, eps2, mode)
        vel = vel + a * (0.5 * dt)
    return pos, vel, a


def leapfrog_rollout(pos, vel, mass, eps2, dt, steps, checkpoint_every=8,
                     mode="fp64"):
    """Kick-drift-kick leapfrog for ``steps`` steps of ``dt``; returns the
    final (pos, vel).  Differentiable in pos, vel and mass; with
    ``checkpoint_every`` > 0 only the segment boundaries are kept for the
    backward pass."""
    a = acc(pos, mass, eps2, mode)
    done = 0
    while done < steps:
        k = min(checkpoint_every or steps, steps - done)
        if checkpoint_every and torch.is_grad_enabled():
            pos, vel, a = checkpoint(_segment, pos, vel, a, mass, eps2, dt, k,
                                     mode, use_reentrant=False)
        else:
            pos, vel, a = _segment(pos, vel, a, mass, eps2, dt, k, mode)
        done += k
    return pos, vel
