"""QuadraticModel (Mandel & Agol 2002) on MI355X.

Drop-in for ``pytransit.models.ma_quadratic.QuadraticModel`` (ma_quadratic.py:37-140):
the constructor, ``set_data`` and ``evaluate(k, ldc, t0, p, a, i, e=None, w=None,
copy=True)`` with the reference's layouts: ``ldc`` is ``[npv, 2*npb]`` (u1, u2 per
passband; a 3-D ``[npv, npb, 2]`` is reshaped), a 1-D ``k`` or ``t0`` is one value
per pv.  The model shares the RoadRunner path's Taylor orbit and contact window
(widened by 0.025 d, ma_quadratic_nb.py:558-561) and evaluates the analytic flux
in FP64 on the GPU.  ``interpolate=True`` is accepted and evaluates the exact
model (the reference's interpolation tables are an approximation of it).
"""
from typing import List, Optional, Union

import numpy as np
import torch
from numpy import isscalar, ndarray

from . import _lib
from ._lib import RRLcData, RRParams, check, ptr, stream_handle
from .rrmodel import _as_dev
from .transitmodel import TransitModel

__all__ = ['QuadraticModel']


class QuadraticModel(TransitModel):
    """Transit model with quadratic limb darkening (Mandel & Agol, ApJ 580, L171-L175 2002)."""

    def __init__(self, interpolate: bool = False, klims: tuple = (0.01, 0.5), nk: int = 256, nz: int = 256,
                 device: Optional[Union[str, torch.device]] = None):
        super().__init__()
        self.interpolate = interpolate
        self.klims, self.nk, self.nz = klims, (nk if interpolate else 0), (nz if interpolate else 0)
        self._device = torch.device(device) if device is not None else None
        self._ws = {}   # (device, stream) -> workspace; see RoadRunnerModel._workspace

    @property
    def device(self) -> torch.device:
        if self._device is None:
            self._device = torch.device('cuda', torch.cuda.current_device())
        return self._device

    def evaluate(self, k: Union[float, ndarray], ldc: Union[ndarray, List], t0: Union[float, ndarray],
                 p: Union[float, ndarray], a: Union[float, ndarray], i: Union[float, ndarray],
                 e: Optional[Union[float, ndarray]] = None, w: Optional[Union[float, ndarray]] = None,
                 copy: bool = True):
        """Evaluate the model for scalar parameters or a population (ma_quadratic.py:74-140).

        Returns the squeezed flux ``[npv, npt]`` (numpy with ``copy=True``, the device
        tensor otherwise).  A NaN radius ratio or semi-major axis gives an inf row.
        """
        flux = self._evaluate_device(k, ldc, t0, p, a, i, e, w).squeeze()
        return _lib.to_host(flux) if copy else flux

    def _evaluate_device(self, k, ldc, t0, p, a, i, e=None, w=None):
        lib = _lib.load()
        device = self.device
        if self.time is None:
            raise ValueError('set_data() must be called before evaluate()')
        if isscalar(p):
            npv = 1
            k = torch.atleast_2d(_as_dev(k, device))
            t0 = torch.atleast_2d(_as_dev(t0, device))
            ldc = _as_dev(ldc, device).reshape(1, -1)
        else:
            npv = np.size(p) if not isinstance(p, torch.Tensor) else p.numel()
            k = _as_dev(k, device)
            t0 = _as_dev(t0, device)
            k = k.reshape(-1, 1) if k.dim() == 1 else torch.atleast_2d(k)
            t0 = t0.reshape(-1, 1) if t0.dim() == 1 else torch.atleast_2d(t0)
            ldc = _as_dev(ldc, device)
            ldc = ldc.reshape(npv, -1) if ldc.dim() == 3 else torch.atleast_2d(ldc)
        if ldc.shape[1] != 2 * self.npb:
            raise ValueError("The quadratic model needs two limb darkening coefficients per passband")
        if self.epids.max() != t0.shape[1] - 1:
            raise ValueError("The number of transit centers must equal to the number of individual epoch IDs.")
        if k.shape[1] not in (1, self.npb):
            raise ValueError('Radius ratios should be given either as an [npv, 1] or [npv, npb] array.')
        prm = []
        for name, x, dflt in (('p', p, None), ('a', a, None), ('i', i, None), ('e', e, 0.0), ('w', w, 0.0)):
            x = _as_dev(dflt if x is None else x, device).reshape(-1)
            if x.numel() not in (1, npv):
                raise ValueError(f"{name} has {x.numel()} values, expected 1 or npv={npv}")
            prm.append(x.expand(npv).contiguous() if x.numel() != npv else x)
        p, a, inc, e, w = prm
        k = k.expand(npv, k.shape[1]).contiguous()
        t0 = t0.expand(npv, t0.shape[1]).contiguous()
        ldc = ldc.expand(npv, ldc.shape[1]).contiguous()
        d = self._device_data(device)
        lc = RRLcData(npt=self.npt, nlc=self.nlc, npb=self.npb, nep=t0.shape[1], sorted=int(d['sorted']),
                      uniform_nsamples=d['uniform_ns'], times=ptr(d['times']), lcids=ptr(d['lcids']),
                      pbids=ptr(d['pbids']), epids=ptr(d['epids']), nsamples=ptr(d['nsamples']),
                      exptimes=ptr(d['exptimes']))
        pv = RRParams(npv=npv, kcols=k.shape[1], t0cols=t0.shape[1], pad_=0, k=ptr(k), t0=ptr(t0), p=ptr(p),
                      a=ptr(a), inc=ptr(inc), e=ptr(e), w=ptr(w))
        nbytes = max(int(lib.ma_workspace_bytes(npv)), 256)
        key = (str(device), stream_handle(device))
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
            self._ws[key] = ws
        flux = torch.empty((npv, self.npt), dtype=torch.float64, device=device)
        check(lib.ma_evaluate(_lib.ctypes.byref(lc), _lib.ctypes.byref(pv), ptr(ldc), ptr(ws),
                              ws.numel(), ptr(flux), stream_handle(device)))
        return flux

    def __call__(self, *args, **kwargs):
        return self.evaluate(*args, **kwargs)
