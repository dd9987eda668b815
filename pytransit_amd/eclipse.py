"""Secondary-eclipse models on MI355X: ``EclipseModel`` and
``EclipseSpectroscopyModel``.

Drop-ins for ``pytransit.models.new_eclipse_model.EclipseModel``
(new_eclipse_model.py:29-70, kernel roadrunner/model_eclipse.py:11-81) and
``pytransit.models.roadrunner.esmodel.EclipseSpectroscopyModel`` (esmodel.py:40-
93, kernel roadrunner/model_ecspec.py:13-65): the same constructors, ``set_data``
and ``evaluate`` arguments.  Both run on the GPU in FP64 (``ec_kernels.hip``):
the Taylor orbit about mid-eclipse, its contact window and the lens area of the
star over the planet.

* ``EclipseModel.evaluate`` returns the visible planet disk area, pi k^2 outside
  the eclipse (model_eclipse.py:73-79), as ``[npv, npt]`` squeezed.  Like the
  reference model it passes the light curves' passband indices where the kernel
  takes epoch indices (new_eclipse_model.py:70), so the columns of ``t0`` are
  indexed by passband.
* ``EclipseSpectroscopyModel.evaluate`` returns ``1 - (f A / pi) / (1 + f k^2)``
  per passband flux ratio f, ``[npv, npb, npt]``, with the first light curve's
  nsamples and exposure time and a 1-D ``t0`` (esmodel.py:84-88).

Where the reference is undefined the behaviour is defined here: scalar or
length-1 ``e``/``w``/``rstar`` are broadcast to npv (the kernels read ``e[ipv]``
past a length-1 array).
"""
from typing import Optional, Union

import numpy as np
import torch
from numpy import ndarray

from . import _lib
from ._lib import RRLcData, RRParams, check, ptr, stream_handle
from .rrmodel import _as_dev
from .transitmodel import TransitModel

__all__ = ['EclipseModel', 'EclipseSpectroscopyModel']


class _EclipseBase(TransitModel):
    def __init__(self, device=None):
        super().__init__()
        self._device = torch.device(device) if device is not None else None
        self._ws = {}

    @property
    def device(self) -> torch.device:
        if self._device is None:
            self._device = torch.device('cuda', torch.cuda.current_device())
        return self._device

    def _workspace(self, nbytes):
        key = (str(self.device), stream_handle(self.device))
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
            self._ws[key] = ws
        return ws

    def _vec(self, name, x, npv, default=None):
        x = _as_dev(default if x is None else x, self.device).reshape(-1)
        if x.numel() not in (1, npv):
            raise ValueError(f"{name} has {x.numel()} values, expected 1 or npv={npv}")
        return x.expand(npv).contiguous() if x.numel() != npv else x


class EclipseModel(_EclipseBase):
    """Secondary eclipse of a uniform planet disk (new_eclipse_model.py:29-70)."""

    def __init__(self, device: Optional[Union[str, torch.device]] = None) -> None:
        super().__init__(device)

    def evaluate(self, k, t0, p, a, i, e=None, w=None, rstar: float = 1.0, copy: bool = True):
        """Visible planet disk area over the secondary eclipse (new_eclipse_model.py:34-70).

        k is a float or ``[npv]``; t0 a float, ``[npv]`` or ``[npv, nep]``; ``rstar``
        the stellar radius in solar radii (for the light travel time).  Invalid
        parameter vectors (a <= 1, a NaN, e < 0) give NaN rows.
        """
        flux = self._evaluate_device(k, t0, p, a, i, e, w, rstar).squeeze()
        return _lib.to_host(flux) if copy else flux

    def __call__(self, *args, **kwargs):
        return self.evaluate(*args, **kwargs)

    def _evaluate_device(self, k, t0, p, a, i, e=None, w=None, rstar=1.0):
        lib = _lib.load()
        if self.time is None:
            raise ValueError('set_data() must be called before evaluate()')
        dev = self.device
        k = _as_dev(k, dev).reshape(-1)
        npv = k.numel()
        t0 = _as_dev(t0, dev).reshape(npv, -1).contiguous()
        # new_eclipse_model.py:70 passes pbids where the kernel takes epoch indices
        if self.pbids.max() >= t0.shape[1]:
            raise ValueError(f"t0 has {t0.shape[1]} columns but the passband indices (the eclipse model's epoch "
                             f"indices) go up to {self.pbids.max()}")
        p, a, inc = (self._vec(n, x, npv) for n, x in (('p', p), ('a', a), ('i', i)))
        e, w = self._vec('e', e, npv, 0.0), self._vec('w', w, npv, 0.0)
        rs = self._vec('rstar', rstar, npv)
        d = self._device_data(dev)
        lc = RRLcData(npt=self.npt, nlc=self.nlc, npb=self.npb, nep=t0.shape[1], sorted=int(d['sorted']),
                      uniform_nsamples=d['uniform_ns'], times=ptr(d['times']), lcids=ptr(d['lcids']),
                      pbids=ptr(d['pbids']), epids=ptr(d['pbids']), nsamples=ptr(d['nsamples']),
                      exptimes=ptr(d['exptimes']))
        prm = RRParams(npv=npv, kcols=1, t0cols=t0.shape[1], pad_=0, k=ptr(k), t0=ptr(t0), p=ptr(p), a=ptr(a),
                       inc=ptr(inc), e=ptr(e), w=ptr(w))
        ws = self._workspace(int(lib.ec_workspace_bytes(npv)))
        flux = torch.empty((npv, self.npt), dtype=torch.float64, device=dev)
        byref = _lib.ctypes.byref
        check(lib.eclipse_evaluate(byref(lc), byref(prm), ptr(rs), ptr(ws), ws.numel(), ptr(flux),
                                   stream_handle(dev)))
        return flux


class EclipseSpectroscopyModel(_EclipseBase):
    """Secondary eclipse observed in many passbands (esmodel.py:40-93)."""

    def __init__(self, parallel: bool = False, device: Optional[Union[str, torch.device]] = None):
        super().__init__(device)
        self.parallel = parallel

    def evaluate(self, f: ndarray, k, t0, p, a, i, e=0.0, w=0.0, rstar=1.0, copy: bool = True):
        """Flux ``[npv, npb, npt]`` for flux ratios ``f`` ``[npv, npb]`` (esmodel.py:46-88).

        k, t0, p, a, i, e, w and rstar are floats or ``[npv]`` vectors.  With
        ``copy=False`` the device tensor is returned.
        """
        flux = self._evaluate_device(f, k, t0, p, a, i, e, w, rstar)
        return _lib.to_host(flux) if copy else flux

    def __call__(self, f, k, t0, p, a, i, e=0.0, w=0.0, rstar=1.0):
        return self.evaluate(f, k, t0, p, a, i, e, w, rstar)

    def _evaluate_device(self, f, k, t0, p, a, i, e=0.0, w=0.0, rstar=1.0):
        lib = _lib.load()
        if self.time is None:
            raise ValueError('set_data() must be called before evaluate()')
        dev = self.device
        k = _as_dev(k, dev).reshape(-1)
        npv = k.numel()
        fr = _as_dev(f, dev)
        if fr.dim() != 2:
            raise ValueError(" The flux ratio must be given as a 2D array with shape (npv, npb)")
        if fr.shape[0] != npv:
            raise ValueError(f"The flux ratio has {fr.shape[0]} rows, the radius ratio {npv} values")
        npb = fr.shape[1]
        t0 = self._vec('t0', t0, npv)
        p, a, inc, e, w, rs = (self._vec(n, x, npv) for n, x in
                               (('p', p), ('a', a), ('i', i), ('e', e), ('w', w), ('rstar', rstar)))
        d = self._device_data(dev)
        lc = RRLcData(npt=self.npt, nlc=self.nlc, npb=self.npb, nep=1, sorted=int(d['sorted']),
                      uniform_nsamples=d['uniform_ns'], times=ptr(d['times']), lcids=ptr(d['lcids']),
                      pbids=ptr(d['pbids']), epids=ptr(d['epids']), nsamples=ptr(d['nsamples']),
                      exptimes=ptr(d['exptimes']))
        prm = RRParams(npv=npv, kcols=1, t0cols=1, pad_=0, k=ptr(k), t0=ptr(t0), p=ptr(p), a=ptr(a), inc=ptr(inc),
                       e=ptr(e), w=ptr(w))
        ws = self._workspace(int(lib.ec_workspace_bytes(npv)))
        flux = torch.empty((npv, npb, self.npt), dtype=torch.float64, device=dev)
        byref = _lib.ctypes.byref
        check(lib.es_evaluate(byref(lc), byref(prm), ptr(rs), ptr(fr), npb, ptr(ws), ws.numel(), ptr(flux),
                              stream_handle(dev)))
        return flux
