"""torch.autograd bridge for the fused log-likelihood and its hand-written adjoint.

``rr_lnlike(theta, model, obs, sigma)`` returns lnL per parameter vector for
``theta[npv, 7] = (k, t0, p, a, i, c0, c1)`` (c0, c1: quadratic or power-2
coefficients) and back-propagates through the adjoint computed by the same
kernel launch (``RoadRunnerModel.lnlike``), so the flux is never materialised.
"""
import torch

__all__ = ['RRLogLikelihood', 'rr_lnlike']


class RRLogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, model, obs, sigma, e=0.0, w=0.0):
        th = theta.detach().to(torch.float64)
        k, t0, p, a, i = th[:, 0], th[:, 1], th[:, 2], th[:, 3], th[:, 4]
        ldc = th[:, 5:7].reshape(-1, 1, 2)
        lnl, g = model._lnlike_device(k, ldc, t0, p, a, i, obs, sigma, e, w, grad=True)
        ctx.save_for_backward(g)
        return lnl.to(theta.dtype)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return (gout.to(g.dtype)[:, None] * g).to(gout.dtype), None, None, None, None, None


def rr_lnlike(theta, model, obs, sigma, e=0.0, w=0.0):
    """Differentiable lnL[npv] of ``theta[npv, 7]`` under ``model`` (a RoadRunnerModel
    with set_data called, 'quadratic' or 'power-2' limb darkening)."""
    return RRLogLikelihood.apply(theta, model, obs, sigma, e, w)
