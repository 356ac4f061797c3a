"""torch.autograd binding of the differentiable flow past a sphere.

``SphereRunFunction.apply(theta, run)`` maps theta = (drift speed [m/s], sphere potential [V],
Te [eV]) -- a float64 tensor of 3 -- to (ion current to the sphere [A], ion number density
after the last step, potential after the last step), the latter two float64 tensors of nn
nodes on the GPU in the reference's flat order U = k*ni*nj + j*ni + i
(ch3/ver2/Field.h:161).  Its backward is the library's hand-written adjoint
(pic_adjoint_backward), called through the same C ABI as everything else; torch supplies
device memory and composes the loss.

``sphere_loss`` is the loss of the differentiable sphere case: the ion current to the sphere
plus the L2 norm of the density in the wake behind it (``SphereAdjointRun.wake_mask``), each
relative to its value at a reference theta.
"""
from __future__ import annotations

import numpy as np
import torch

import pic_amd as P


class SphereRunFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, run: P.SphereAdjointRun):
        dev = torch.device("cuda", torch.cuda.current_device())
        den = torch.empty(run.nn, dtype=torch.float64, device=dev)
        phi = torch.empty(run.nn, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        cur = run.forward(theta.detach().cpu().numpy(), den.data_ptr(), phi.data_ptr())
        ctx.run = run
        ctx.theta = theta.detach().cpu().numpy().copy()
        return (torch.tensor(cur, dtype=torch.float64, device=theta.device), den, phi)

    @staticmethod
    def backward(ctx, g_cur, g_den, g_phi):
        run = ctx.run
        if run.theta is None or not np.array_equal(run.theta, ctx.theta):
            run.forward(ctx.theta)  # the run holds another theta's checkpoints: recompute
        dden = g_den.contiguous() if g_den is not None else None
        dphi = g_phi.contiguous() if g_phi is not None else None
        torch.cuda.synchronize()  # torch's kernels that wrote the gradients have finished
        gc = float(g_cur) if g_cur is not None else 0.0
        grad = run.backward(gc, dden.data_ptr() if dden is not None else 0,
                            dphi.data_ptr() if dphi is not None else 0)
        return torch.tensor(grad, dtype=torch.float64), None


def sphere_loss(cur, den, mask, cur_ref, den_ref):
    """current / current_ref + mean(den^2 over the wake) / den_ref (torch ops)."""
    return cur / cur_ref + (den[mask] ** 2).mean() / den_ref
