"""torch.autograd binding of the adjoint misfit gradient (sfdx_adjoint).

    phi = mt_misfit(log_scale, model, obs)        # scalar tensor
    phi.backward()                                # log_scale.grad = d phi / d log_scale

`log_scale` (float64, one entry per cell) multiplies every cell's conductivity
tensor by exp(log_scale); for an isotropic model with log_scale = log(sigma) -
log(sigma0) this is the gradient with respect to log sigma.  The forward pass
runs, per frequency, the forward solve, the misfit and the adjoint solve on
the GPU and keeps only the gradient (ncell doubles): one frequency's fields are
resident at a time, so HBM use does not grow with the number of frequencies.
"""
from __future__ import annotations

import numpy as np
import torch

from . import sfdx


class _MTMisfit(torch.autograd.Function):
    @staticmethod
    def forward(fctx, log_scale, model, obs, opts, context, rd):
        ls = log_scale.detach().to("cpu", torch.float64).numpy()
        sigma = np.asarray(model.sigma, np.float64) * np.exp(ls)[:, None]
        phi, grad, _ = context.adjoint(model.a, model.b, model.c, model.nair, sigma, model.freqs, model.recv, obs,
                                       opts, rd=rd)
        fctx.save_for_backward(torch.from_numpy(grad).to(log_scale.device))
        return torch.tensor(phi, dtype=torch.float64, device=log_scale.device)

    @staticmethod
    def backward(fctx, gout):
        (grad,) = fctx.saved_tensors
        return gout * grad, None, None, None, None, None


_default_ctx = None


def mt_misfit(log_scale: torch.Tensor, model, obs, opts=None, context: sfdx.Context | None = None, rd: int = 2):
    """Misfit of `model` (a, b, c, nair, sigma, freqs, recv) with every cell's
    tensor scaled by exp(log_scale) against obs rows (rho_xy, rho_yx, phi_xy,
    phi_yx), differentiable with respect to log_scale."""
    global _default_ctx
    if context is None:
        if _default_ctx is None:
            _default_ctx = sfdx.Context(0)
        context = _default_ctx
    return _MTMisfit.apply(log_scale, model, obs, opts, context, rd)
