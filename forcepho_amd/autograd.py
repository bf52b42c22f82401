"""torch.autograd.Function around the hand-written adjoint.

``LnLike.apply(theta, batch, splines)`` returns lnlike = -chi2/2 per patch
(forcepho/model.py:331) for the parameter vector ``theta`` (float64, on the
device; per source [flux_1..flux_NB, ra, dec, q, pa, nsersic, rh], patches
concatenated).  The forward pass already computes the exact gradient with the
HIP kernels (the adjoint of the mixture model: sum over pixels of
-2 r ierr^2 dm/dtheta, common_kernel.cc:95-143, stacked as
forcepho/model.py:370-387); backward only scales it by the incoming
gradient.  This replaces the reference's Theano ops (forcepho/model.py:
515-551).
"""

import torch

__all__ = ["LnLike", "lnlike"]


class LnLike(torch.autograd.Function):

    @staticmethod
    def forward(ctx, theta, batch, splines):
        ll, grad = batch.lnlike_and_grad(theta, splines)
        ctx.save_for_backward(grad.clone())
        ctx.param_patch = batch.param_patch
        return ll.clone()

    @staticmethod
    def backward(ctx, g):
        grad, = ctx.saved_tensors
        return g[ctx.param_patch] * grad, None, None


def lnlike(theta, batch, splines):
    """lnlike per patch with autograd support."""
    return LnLike.apply(theta, batch, splines)
