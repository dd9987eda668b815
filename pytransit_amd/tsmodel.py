"""TransmissionSpectroscopyModel (TSModel) on MI355X.

Drop-in for ``pytransit.models.roadrunner.tsmodel.TransmissionSpectroscopyModel``
(``tsmodel.py:44-136``): the RoadRunnerModel constructor and ``set_data``, and
``evaluate(k, ldc, t0, p, a, i, e=0.0, w=0.0, copy=True)`` returning the flux
``[npv, npb, npt]``.  k is ``[npv, npb]`` (one radius ratio per passband) and the
number of passbands comes from ``ldc.shape[1]`` (tsmodel.py:101); ``nsamples[0]``
and ``exptimes[0]`` apply to every point and ``lcids``/``pbids`` are not used
(model_trspec.py:67-68,79-80).  With ``precompute_weights=False`` (the default)
each parameter vector gets exact 2-D weights at mean(k); with True the 3-D table
is used (tsmodel.py:117-120).  The arithmetic is the serial kernel's
(model_trspec.py:11-93); the reference's parallel variant is not reproduced.
"""
from typing import List, Union

import numpy as np
import torch
from numpy import ndarray

from . import _lib
from ._lib import LD_LAW_IDS, RRLcData, RRLd, RRParams, RRTable, check, ptr, stream_handle
from .ldmodel import LDModel
from .rrmodel import RoadRunnerModel, _PRECISIONS, _as_dev, _shape_of

__all__ = ['TransmissionSpectroscopyModel', 'TSModel']


class TransmissionSpectroscopyModel(RoadRunnerModel):

    def evaluate(self, k: Union[float, ndarray], ldc: Union[ndarray, List],
                 t0: Union[float, ndarray], p: Union[float, ndarray], a: Union[float, ndarray],
                 i: Union[float, ndarray], e: Union[float, ndarray] = 0.0, w: Union[float, ndarray] = 0.0,
                 copy: bool = True):
        """Evaluate the model for a set of parameter vectors (tsmodel.py:46-130).

        Returns the flux ``[npv, npb, npt]`` as numpy (``copy=True``) or as the device
        tensor (``copy=False``).  Invalid parameter vectors (a <= 1, a NaN, e < 0)
        give NaN (model_trspec.py:37-39).
        """
        flux = self._evaluate_device(k, ldc, t0, p, a, i, e, w)
        return _lib.to_host(flux) if copy else flux

    def _evaluate_device(self, k, ldc, t0, p, a, i, e=0.0, w=0.0, flux_events=None):
        lib = _lib.load()
        device = self.device
        if self.time is None:
            raise ValueError('set_data() must be called before evaluate()')
        k = torch.atleast_2d(_as_dev(k, device))
        npv, npb = k.shape
        ldc_shape = _shape_of(ldc)
        if npv > 1 and len(ldc_shape) != 3 and not isinstance(self.ldmodel, LDModel):
            raise ValueError("The limb darkening parameters (ldp) should be given as a 3D array with shape "
                             "[npv, npb, nldp] when evaluating the model for a set of parameters (npv > 1).")
        prm_t = []
        for name, x in (('t0', t0), ('p', p), ('a', a), ('i', i), ('e', e), ('w', w)):
            x = _as_dev(x, device).reshape(-1)
            if x.numel() not in (1, npv):
                raise ValueError(f"{name} has {x.numel()} values, expected 1 or npv={npv}")
            prm_t.append(x.expand(npv).contiguous() if x.numel() != npv else x)
        t0, p, a, inc, e, w = prm_t
        self.npb = npb

        keep = []
        if self.ldname is not None:
            law, ncoef = LD_LAW_IDS[self.ldname]
            c = _as_dev(ldc, device)
            if c.dim() == 2:
                c = c.reshape(1, c.shape[0], c.shape[1])
            if c.dim() != 3:
                raise ValueError('ldc must be an [npv, npb, nldc] array')
            if c.shape[1] != npb:
                raise ValueError("The transmission spectrum transit model requires that the number or radius "
                                 "ratios and the number of passbands match.")
            if c.shape[2] < ncoef:
                raise ValueError(f"The '{self.ldname}' law needs {ncoef} coefficients per passband")
            c = c.expand(npv, npb, c.shape[2]).contiguous()
            keep.append(c)
            ld = RRLd(law=law, nldc=c.shape[2], ldc=ptr(c), ldp=0, istar=0)
        else:
            if isinstance(self.ldmodel, LDModel):
                ldp, istar = self.ldmodel(self.mu, ldc)
                ldp = _as_dev(ldp, device)
                if istar is None:
                    istar = torch.empty(ldp.shape[:2], dtype=torch.float64, device=device)
                    mu_d = torch.as_tensor(self.mu, device=device)
                    check(lib.rr_integrate_profiles(ptr(mu_d), self.mu.size, ptr(ldp), ldp.shape[0], ldp.shape[1],
                                                    ptr(istar), stream_handle(device)))
                else:
                    istar = _as_dev(istar, device)
            else:
                if isinstance(ldc, torch.Tensor):
                    ldc = ldc.detach().cpu().numpy()
                ldp, istar = self._host_ld(np.asarray(ldc), npb)
                ldp, istar = _as_dev(ldp, device), _as_dev(istar, device)
            if ldp.dim() != 3 or ldp.shape[1] != npb:
                raise ValueError("The transmission spectrum transit model requires that the number or radius "
                                 "ratios and the number of passbands match.")
            ldp, istar = self._broadcast_ld(ldp, istar.reshape(ldp.shape[0], ldp.shape[1]), npv, npb)
            keep += [ldp, istar]
            ld = RRLd(law=_lib.RR_LD_TABULATED, nldc=0, ldc=0, ldp=ptr(ldp), istar=ptr(istar))

        tab, _k = self._table(device)
        if not self.interpolate:
            tab = RRTable(nk=tab.nk, ng=tab.ng, nz=tab.nz, pad_=0, kmin=tab.kmin, kmax=tab.kmax, dk=tab.dk, dg=tab.dg,
                          weights=0, weights32=0, ze=tab.ze, mu=tab.mu)
        d = self._device_data(device)
        lc = RRLcData(npt=self.npt, nlc=self.nlc, npb=npb, nep=1, sorted=int(d['sorted']),
                      uniform_nsamples=int(self._per_lc(self.nsamples, np.int64)[0]), times=ptr(d['times']),
                      lcids=ptr(d['lcids']), pbids=ptr(d['pbids']), epids=ptr(d['epids']),
                      nsamples=ptr(d['nsamples']), exptimes=ptr(d['exptimes']))
        prm = RRParams(npv=npv, kcols=npb, t0cols=1, pad_=0, k=ptr(k), t0=ptr(t0), p=ptr(p), a=ptr(a), inc=ptr(inc),
                       e=ptr(e), w=ptr(w))
        prec = _PRECISIONS[self.precision]
        dtype = torch.float64 if prec == _lib.RR_FP64 else torch.float32
        flux = torch.empty((npv, npb, self.npt), dtype=dtype, device=device)
        nbytes = lib.ts_workspace_bytes(npv, npb, self.ng, self.mu.size, prec)
        ws = self._workspace(device, nbytes)
        if flux_events is not None:
            flux_events[0].record()
        check(lib.ts_evaluate(prec, _lib.ctypes.byref(tab), _lib.ctypes.byref(lc), _lib.ctypes.byref(prm),
                              _lib.ctypes.byref(ld), ptr(ws), ws.numel(), ptr(flux), stream_handle(device)))
        if flux_events is not None:
            flux_events[1].record()
        return flux

    def __call__(self, k, ldc, t0, p, a, i, e=0.0, w=0.0, copy: bool = True):
        return self.evaluate(k, ldc, t0, p, a, i, e, w, copy)


TSModel = TransmissionSpectroscopyModel
