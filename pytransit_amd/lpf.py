"""RoadRunner with the argument layouts of PyTransit's log-posterior functions.

``BaseLPF.transit_model`` (pytransit/lpf/lpf.py:435-443) calls its model as
``tm.evaluate(k[npv, 1], ldc[npv, 2*npb], t0[npv], p, a, i)``: the limb-darkening
coefficients are the quadratic (u, v) pairs of every passband side by side (the
output of ``map_ldc``, lpf.py:85-91) and the zero epoch is one value per pv.  The
reference RoadRunnerModel reads those layouts wrongly (a 2-D ldc is reshaped as
(1, npb', nldc) and a 1-D t0 becomes a row).  ``RoadRunnerLPFModel`` accepts them,
so it can be set as an LPF's ``tm``, and adds the fused likelihood of
``BaseLPF.lnlikelihood`` with white noise (lpf.py:454-475,
loglikelihood/wnloglikelihood.py:22-35).
"""
import numpy as np
import torch

from .rrmodel import RoadRunnerModel

__all__ = ['RoadRunnerLPFModel', 'map_ldc']


def map_ldc(ldc):
    """Kipping's (q1, q2) to quadratic (u, v) per passband, as lpf.py:85-91:
    u = 2 sqrt(q1) q2, v = sqrt(q1) (1 - 2 q2), columns interleaved."""
    ldc = np.atleast_2d(np.asarray(ldc, float))
    a, b = np.sqrt(ldc[:, 0::2]), 2.0 * ldc[:, 1::2]
    uv = np.zeros_like(ldc)
    uv[:, 0::2] = a * b
    uv[:, 1::2] = a * (1.0 - b)
    return uv


def _pairs(ldc, npb):
    """[npv, 2*npb] (or [2*npb]) interleaved pairs -> [npv, npb, 2]."""
    if isinstance(ldc, torch.Tensor):
        c = ldc if ldc.dim() == 2 else ldc.reshape(1, -1)
    else:
        c = np.atleast_2d(np.asarray(ldc, float))
    if c.shape[1] != 2 * npb:
        raise ValueError("The quadratic model needs two limb darkening coefficients per passband")
    return c.reshape(c.shape[0], npb, 2)


class RoadRunnerLPFModel(RoadRunnerModel):
    """RoadRunnerModel('quadratic') taking BaseLPF's layouts: ``ldc`` as
    [npv, 2*npb] interleaved (u, v) pairs, ``t0`` as [npv] (or [npv, nep])."""

    def __init__(self, **kwargs):
        kwargs.pop('ldmodel', None)
        super().__init__('quadratic', **kwargs)

    def _evaluate_device(self, k, ldc, t0, p, a, i, e=0.0, w=0.0, flux_events=None):
        return super()._evaluate_device(k, _pairs(ldc, self.npb), t0, p, a, i, e, w, flux_events)

    def lnlikelihood(self, k, ldc, t0, p, a, i, obs, sigma, e=0.0, w=0.0, copy=True):
        """Log likelihood per pv of ``obs`` (one light curve) with white noise
        ``sigma``; the flux is never written to memory."""
        return self.lnlike(k, _pairs(ldc, 1), t0, p, a, i, obs, sigma, e, w, grad=False, copy=copy)
