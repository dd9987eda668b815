"""RoadRunnerModel on MI355X.

Drop-in for ``pytransit.models.roadrunner.rrmodel.RoadRunnerModel``
(``rrmodel.py:47-238``): the same constructor arguments, ``set_data`` and
``evaluate(k, ldc, t0, p, a, i, e=0.0, w=0.0, copy=True)``.  Two keywords are
added: ``precision`` ('fp64', the reference's arithmetic, or 'mixed': FP64 phase
and setup, FP32 flux) and ``device``.

The weight table (``calculate_weights_3d``, built in the reference's ``__init__``)
is built by a HIP kernel the first time the model is used on a device, and stays
resident there.  ``evaluate`` runs three kernels on torch's current stream:
limb-darkening profiles, per-parameter-vector setup (Taylor orbit, contact window,
limb-darkening means) and the flux.  With ``copy=True`` it returns a numpy array,
with ``copy=False`` the device tensor (no host synchronisation).

Shapes where the reference is undefined are defined here:
* a 1-D ``k`` or ``t0`` of length npv > 1 is read as ``(npv, 1)`` (the reference
  turns it into ``(1, npv)`` and reads out of bounds, Q4);
* scalar ``e``/``w`` (the defaults) are broadcast to npv (the reference reads
  ``e[ipv]`` past a length-1 array for npv > 1);
* ``ldc`` is ``(npv|1, npb|1, nldc)`` after the reference's reshape rules
  (numba/ldmodels.py:144-150); size-1 axes are broadcast.
"""
from typing import Callable, List, Optional, Tuple, Union
from warnings import warn

import numpy as np
import torch
from numpy import isscalar, linspace, ndarray, pi, sqrt, zeros
from scipy.integrate import trapezoid

from . import _lib
from ._lib import LD_LAW_IDS, RRLcData, RRLd, RRParams, RRTable, check, ptr, stream_handle
from .ldmodel import LDModel
from .transitmodel import TransitModel

__all__ = ['RoadRunnerModel']

_PRECISIONS = {'fp64': _lib.RR_FP64, 'mixed': _lib.RR_MIXED}


def _as_dev(x, device, dtype=torch.float64) -> torch.Tensor:
    """Contiguous device tensor of ``dtype`` (no copy if ``x`` already is one)."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(x, dtype=np.float64)), device=device).to(dtype)


def _shape_of(x):
    return tuple(x.shape) if isinstance(x, (torch.Tensor, np.ndarray)) else np.shape(x)


def _size_of(x):
    s = _shape_of(x)
    return int(np.prod(s)) if len(s) else 1


def _ldc3_shape(shape):
    """Shape after the evaluate_ld reshape rules (numba/ldmodels.py:144-150)."""
    if len(shape) == 1:
        return (1, 1, shape[0])
    if len(shape) == 2:
        return (1, shape[1], (shape[0] * shape[1]) // shape[1] if shape[1] else 0)
    return tuple(shape)


class RoadRunnerModel(TransitModel):
    # String keys of the reference (rrmodel.py:48-58); the laws run on the GPU.
    ldmodels = {name: name for name in LD_LAW_IDS}

    def __init__(self, ldmodel: Union[str, Callable, Tuple[Callable, Callable], LDModel] = 'quadratic',
                 precompute_weights: bool = False, klims: tuple = (0.005, 0.5), nk: int = 256,
                 nzin: int = 20, nzlimb: int = 20, zcut: float = 0.7, ng: int = 100,
                 nthreads: int = 1, small_planet_limit: float = 0.05,
                 precision: str = 'fp64', device: Optional[Union[str, torch.device]] = None, **kwargs):
        """The RoadRunner transit model by Parviainen (2020); rrmodel.py:60-154.

        Parameters are the reference's; ``nthreads`` is accepted and ignored (the
        model runs on the GPU).  ``precision`` is 'fp64' or 'mixed'; ``device`` is a
        torch device (default: the current CUDA/HIP device).
        """
        super().__init__()
        if 'interpolate' in kwargs:
            warn("The 'interpolate' argument has been replaced by 'precompute_weights' and will be removed in "
                 "the future.", FutureWarning)
            self.interpolate: bool = kwargs.get('interpolate', False)
        else:
            self.interpolate: bool = precompute_weights
        if 'parallel' in kwargs:
            warn("The 'parallel' argument has been replaced by 'nthreads' and will be removed in the future.",
                 FutureWarning)
        self.nthreads = nthreads
        self.parallel = nthreads > 1
        self.splimit = small_planet_limit

        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {list(_PRECISIONS)}")
        self.precision = precision
        self._device = torch.device(device) if device is not None else None

        # Limb darkening model (rrmodel.py:107-131)
        self.ldname: Optional[str] = None
        if isinstance(ldmodel, str):
            if ldmodel not in self.ldmodels:
                print(f"Unknown limb darkening model: {ldmodel}. Choose from [{', '.join(self.ldmodels.keys())}] "
                      f"or supply a callable function.")
                raise KeyError(ldmodel)
            self.ldname = ldmodel
            self.ldmodel = ldmodel
            self.ldmmean = None
        elif isinstance(ldmodel, LDModel):
            self.ldmodel = ldmodel
            self.ldmmean = ldmodel._integrate
        elif callable(ldmodel):
            self.ldmodel = ldmodel
            self.ldmmean = None
        elif isinstance(ldmodel, tuple) and callable(ldmodel[0]) and callable(ldmodel[1]):
            self.ldmodel = ldmodel[0]
            self.ldmmean = ldmodel[1]
        else:
            raise NotImplementedError

        self.klims = klims
        self._ldmu = linspace(1, 0, 200)
        self._ldz = sqrt(1 - self._ldmu ** 2)
        self._tables = {}
        self._ws = {}
        self.init_integration(nzin, nzlimb, zcut, ng, nk)

    # ---------------------------------------------------------------------------------
    def init_integration(self, nzin, nzlimb, zcut, ng, nk):
        """z grid and weight-table geometry (rrmodel.py:165-173).  The table itself is
        built on the device by :meth:`_table`."""
        if nk < 2 or ng < 2:
            raise ValueError('nk and ng must be >= 2')
        self.nk, self.ng, self.nzin, self.nzlimb, self.zcut = nk, ng, nzin, nzlimb, zcut
        self.ze, self.zm, self.mu = _lib.z_grid(zcut, nzin, nzlimb)
        kmin, kmax = self.klims
        self.dk = (kmax - kmin) / nk
        self.dg = float(np.linspace(0., 1. - 1e-7, ng)[1])
        self._tables = {}

    @property
    def device(self) -> torch.device:
        if self._device is None:
            self._device = torch.device('cuda', torch.cuda.current_device())
        return self._device

    def _table(self, device):
        """Device-resident z grid, mu nodes and weight table (FP64 and FP32 copies, in
        the device layout [nk, nz, ng]); returns (RRTable, keep-alive dict)."""
        key = str(device)
        t = self._tables.get(key)
        if t is None:
            lib = _lib.load()
            ze = torch.as_tensor(self.ze, device=device)
            mu = torch.as_tensor(self.mu, device=device)
            nz = self.ze.size
            W = torch.empty((self.nk, nz, self.ng), dtype=torch.float64, device=device)
            W32 = torch.empty((self.nk, nz, self.ng), dtype=torch.float32, device=device)
            dk = _lib.c_double()
            dg = _lib.c_double()
            check(lib.rr_weights_3d(self.nk, float(self.klims[0]), float(self.klims[1]), ptr(ze), nz, self.ng,
                                    ptr(W), ptr(W32), _lib.ctypes.byref(dk), _lib.ctypes.byref(dg),
                                    stream_handle(device)))
            self.dk, self.dg = dk.value, dg.value
            tab = RRTable(nk=self.nk, ng=self.ng, nz=nz, pad_=0, kmin=float(self.klims[0]),
                          kmax=float(self.klims[1]), dk=self.dk, dg=self.dg, weights=ptr(W), weights32=ptr(W32),
                          ze=ptr(ze), mu=ptr(mu))
            t = (tab, dict(ze=ze, mu=mu, W=W, W32=W32))
            self._tables[key] = t
        return t

    @property
    def weights(self) -> ndarray:
        """The FP64 weight table in the reference's layout W[nk, ng, nz] (a host copy
        of the device table)."""
        return self._table(self.device)[1]['W'].permute(0, 2, 1).contiguous().cpu().numpy()

    def set_data(self, time: Union[ndarray, List], lcids: Optional[Union[ndarray, List]] = None,
                 pbids: Optional[Union[ndarray, List]] = None, nsamples: Optional[Union[ndarray, List]] = None,
                 exptimes: Optional[Union[ndarray, List]] = None,
                 epids: Optional[Union[ndarray, List]] = None) -> None:
        """rrmodel.py:156-163."""
        super().set_data(time, lcids, pbids, nsamples, exptimes, epids)
        self.nep = np.unique(self.epids).size

    # ---------------------------------------------------------------------------------
    def _workspace(self, device, nbytes: int) -> torch.Tensor:
        """The per-pv workspace of torch's current stream on ``device``.  One buffer
        per (device, stream): calls issued on different streams may overlap on the
        device, and the setup of one would overwrite the records the flux kernel of
        the other is still reading.  A buffer is allocated on, and only ever used
        by, its own stream, so the caching allocator can reuse a replaced buffer's
        memory only after that stream's earlier kernels."""
        key = (str(device), stream_handle(device))
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
            self._ws[key] = ws
        return ws

    def _host_ld(self, ldc: ndarray, npb: int):
        """Profiles of a user-supplied (host) law; rrmodel.py:217-227 with
        evaluate_ld/evaluate_ldi's reshape rules (numba/ldmodels.py:142-175)."""
        ldc = np.asarray(ldc, float)
        c3 = ldc.reshape(_ldc3_shape(ldc.shape))
        npv_l, npb_l = c3.shape[:2]
        ldp = zeros((npv_l, npb_l, self.mu.size))
        istar = zeros((npv_l, npb_l))
        for ipv in range(npv_l):
            for ipb in range(npb_l):
                ldp[ipv, ipb] = self.ldmodel(self.mu, c3[ipv, ipb])
                if self.ldmmean is not None:
                    istar[ipv, ipb] = self.ldmmean(c3[ipv, ipb])
                else:
                    prof = self.ldmodel(self._ldmu, c3[ipv, ipb])
                    istar[ipv, ipb] = 2 * pi * trapezoid(self._ldz * prof, self._ldz)
        return ldp, istar

    def _ld_descriptor(self, ldc, npv: int, npb: int, device):
        """Limb darkening of one evaluate call as an RRLd descriptor and the device
        tensors it points to.  Named laws are evaluated by the setup kernel; LDModel
        plug-ins and host callables give tabulated profiles."""
        lib = _lib.load()
        nz = self.mu.size
        if self.ldname is not None:
            law, ncoef = LD_LAW_IDS[self.ldname]
            shp = _ldc3_shape(_shape_of(ldc))
            if shp[2] < ncoef:
                raise ValueError(f"The '{self.ldname}' law needs {ncoef} coefficients per passband, got {shp[2]}")
            c3 = _as_dev(ldc, device).reshape(shp)
            if shp[0] not in (1, npv) or shp[1] not in (1, npb):
                raise ValueError(f"The limb darkening coefficients have shape {shp} after reshaping; expected "
                                 f"[npv={npv}|1, npb={npb}|1, nldc].")
            if shp[0] != npv or shp[1] != npb:
                c3 = c3.expand(npv, npb, shp[2]).contiguous()
            return RRLd(law=law, nldc=shp[2], ldc=ptr(c3), ldp=0, istar=0), (c3,)
        if isinstance(self.ldmodel, LDModel):
            ldp, istar = self.ldmodel(self.mu, ldc)
            ldp = _as_dev(ldp, device)
            if ldp.dim() == 1:
                ldp = ldp.reshape(1, 1, -1)
            if istar is None:
                istar = torch.empty(ldp.shape[:2], dtype=torch.float64, device=device)
                mu_d = torch.as_tensor(self.mu, device=device)
                check(lib.rr_integrate_profiles(ptr(mu_d), nz, ptr(ldp), ldp.shape[0], ldp.shape[1], ptr(istar),
                                                stream_handle(device)))
            else:
                istar = _as_dev(istar, device).reshape(ldp.shape[:2])
        else:
            if isinstance(ldc, torch.Tensor):
                ldc = ldc.detach().cpu().numpy()
            ldp, istar = self._host_ld(ldc, npb)
            ldp, istar = _as_dev(ldp, device), _as_dev(istar, device)
        if ldp.shape[-1] != nz:
            raise ValueError(f"The limb darkening profiles have {ldp.shape[-1]} mu nodes, the model has {nz}")
        ldp, istar = self._broadcast_ld(ldp, istar, npv, npb)
        return RRLd(law=_lib.RR_LD_TABULATED, nldc=0, ldc=0, ldp=ptr(ldp), istar=ptr(istar)), (ldp, istar)

    @staticmethod
    def _broadcast_ld(ldp, istar, npv, npb):
        nv, nb = ldp.shape[0], ldp.shape[1]
        if nv not in (1, npv) or nb not in (1, npb):
            raise ValueError(f"The limb darkening coefficients give profiles for {nv} parameter vectors and {nb} "
                             f"passbands, but the model has npv={npv} and npb={npb}; give ldc as an "
                             f"[npv, npb, nldc] array.")
        if nv != npv or nb != npb:
            ldp = ldp.expand(npv, npb, ldp.shape[2]).contiguous()
            istar = istar.reshape(nv, nb).expand(npv, npb).contiguous()
        return ldp, istar

    def _params(self, k, t0, p, a, i, e, w, device):
        """Normalised FP64 device parameter tensors (rrmodel.py:212,229-230)."""
        npv = 1 if isscalar(p) else _size_of(p)
        ks = _shape_of(k)
        t0s = _shape_of(t0)
        k = _as_dev(k, device)
        t0 = _as_dev(t0, device)
        # Q4 (defined): a 1-D k or t0 of length npv > 1 is one value per pv.
        k = k.reshape(npv, 1) if (len(ks) == 1 and npv > 1 and ks[0] == npv) else torch.atleast_2d(k)
        t0 = t0.reshape(npv, 1) if (len(t0s) == 1 and npv > 1 and t0s[0] == npv) else torch.atleast_2d(t0)
        if k.shape[0] not in (1, npv) or t0.shape[0] not in (1, npv):
            raise ValueError(f"k and t0 need 1 or npv={npv} rows, got {tuple(k.shape)} and {tuple(t0.shape)}")
        k = k.expand(npv, k.shape[1]).contiguous()
        t0 = t0.expand(npv, t0.shape[1]).contiguous()
        out = []
        for name, x in (('p', p), ('a', a), ('i', i), ('e', e), ('w', w)):
            x = _as_dev(x, device).reshape(-1)
            if x.numel() not in (1, npv):
                raise ValueError(f"{name} has {x.numel()} values, expected 1 or npv={npv}")
            out.append(x.expand(npv).contiguous() if x.numel() != npv else x)
        return npv, k, t0, out

    def evaluate(self, k: Union[float, ndarray], ldc: Union[ndarray, List],
                 t0: Union[float, ndarray], p: Union[float, ndarray], a: Union[float, ndarray],
                 i: Union[float, ndarray], e: Union[float, ndarray] = 0.0, w: Union[float, ndarray] = 0.0,
                 copy: bool = True):
        """Evaluate the transit model for scalar or vector parameters (rrmodel.py:175-238).

        Returns the squeezed flux ``[npv, npt]``: a numpy array with ``copy=True``,
        the device tensor with ``copy=False``.  FP64 for precision='fp64', FP32 for
        'mixed'.  Invalid parameter vectors (a <= 1, a NaN, e < 0 or NaN
        limb darkening) give NaN rows (model_full.py:40-42,80-82).
        """
        flux = self._evaluate_device(k, ldc, t0, p, a, i, e, w).squeeze()
        return _lib.to_host(flux) if copy else flux

    def _descriptors(self, k, ldc, t0, p, a, i, e, w):
        """The C-ABI descriptors of one evaluate call: (npv, RRTable, RRLcData,
        RRParams, RRLd, keep), where ``keep`` holds the device tensors they point to."""
        device = self.device
        if self.time is None:
            raise ValueError('set_data() must be called before evaluate()')
        npv, k, t0, (p, a, inc, e, w) = self._params(k, t0, p, a, i, e, w, device)
        npb = self.npb
        if k.shape[1] > 1 and k.shape[1] != npb:
            raise ValueError('Radius ratios should be given either as an [npv, 1] or [npv, npb] array.')
        if self.epids.max() >= t0.shape[1]:
            raise ValueError(f"t0 has {t0.shape[1]} columns but the epoch indices go up to {self.epids.max()}")
        if not isscalar(ldc) and not isinstance(ldc, torch.Tensor):
            ldc = np.atleast_1d(ldc)
        ld, _ld_keep = self._ld_descriptor(ldc, npv, npb, device)
        tab, _keep = self._table(device)
        d = self._device_data(device)
        lc = RRLcData(npt=self.npt, nlc=self.nlc, npb=npb, nep=t0.shape[1], sorted=int(d['sorted']),
                      uniform_nsamples=d['uniform_ns'], times=ptr(d['times']),
                      lcids=ptr(d['lcids']), pbids=ptr(d['pbids']), epids=ptr(d['epids']),
                      nsamples=ptr(d['nsamples']), exptimes=ptr(d['exptimes']))
        prm = RRParams(npv=npv, kcols=k.shape[1], t0cols=t0.shape[1], pad_=0, k=ptr(k), t0=ptr(t0), p=ptr(p),
                       a=ptr(a), inc=ptr(inc), e=ptr(e), w=ptr(w))
        return npv, tab, lc, prm, ld, (k, t0, p, a, inc, e, w, _ld_keep, _keep)

    def _evaluate_device(self, k, ldc, t0, p, a, i, e, w, flux_events=None):
        """All three stages on torch's current stream; returns the [npv, npt] device
        flux.  ``flux_events`` = (start, end) torch.cuda.Event pair recorded around
        the flux kernel alone."""
        lib = _lib.load()
        device = self.device
        npv, tab, lc, prm, ld, _keep = self._descriptors(k, ldc, t0, p, a, i, e, w)
        npb = self.npb
        prec = _PRECISIONS[self.precision]
        dtype = torch.float64 if prec == _lib.RR_FP64 else torch.float32
        flux = torch.empty((npv, self.npt), dtype=dtype, device=device)
        nbytes = lib.rr_workspace_bytes(npv, self.npt, npb, self.ng, self.nlc)
        ws = self._workspace(device, nbytes)
        stream = stream_handle(device)
        byref = _lib.ctypes.byref
        check(lib.rr_setup(prec, byref(tab), byref(lc), byref(prm), byref(ld), ptr(ws), ws.numel(), stream))
        if flux_events is not None:
            flux_events[0].record()
        check(lib.rr_flux(prec, byref(tab), byref(lc), byref(prm), ptr(ws), ws.numel(), ptr(flux), stream))
        if flux_events is not None:
            flux_events[1].record()
        return flux

    def __call__(self, *args, **kwargs):
        return self.evaluate(*args, **kwargs)

    # ---------------------------------------------------------------------------------
    def lnlike(self, k, ldc, t0, p, a, i, obs, sigma, e=0.0, w=0.0, grad: bool = False, copy: bool = True):
        """Gaussian log-likelihood of ``obs`` under the model, per parameter vector,
        fused with the model on the GPU (no flux array is written):

            lnL = sum_t -ln sigma - 1/2 ln(2 pi) - 1/2 ((obs_t - model_t) / sigma)^2

        (the form of lnlike_normal, pytransit/lpf/loglikelihood/wnloglikelihood.py:
        22-35).  With ``grad=True`` also returns d lnL / d(k, t0, p, a, i, c0, c1)
        as ``[npv, 7]`` from the hand-written adjoint, where c0, c1 are the two
        coefficients of the 'quadratic' or 'power-2' law.  One light curve, k of
        shape (npv,) or (npv, 1) inside ``klims``; ``sigma`` is a scalar or one
        value per pv.
        """
        lnl, g = self._lnlike_device(k, ldc, t0, p, a, i, obs, sigma, e, w, grad)
        if copy:
            return (lnl.cpu().numpy(), g.cpu().numpy()) if grad else lnl.cpu().numpy()
        return (lnl, g) if grad else lnl

    def _lnlike_device(self, k, ldc, t0, p, a, i, obs, sigma, e=0.0, w=0.0, grad=True):
        lib = _lib.load()
        device = self.device
        if self.time is None:
            raise ValueError('set_data() must be called before lnlike()')
        if self.nlc != 1:
            raise ValueError('lnlike supports one light curve')
        if self.ldname not in ('quadratic', 'power-2'):
            raise ValueError("lnlike supports the 'quadratic' and 'power-2' laws")
        npv, k, t0, (p, a, inc, e, w) = self._params(k, t0, p, a, i, e, w, device)
        if k.shape[1] != 1:
            raise ValueError('lnlike needs one radius ratio per parameter vector')
        if self.epids.max() >= t0.shape[1]:
            raise ValueError(f"t0 has {t0.shape[1]} columns but the epoch indices go up to {self.epids.max()}")
        ld, _keep = self._ld_descriptor(ldc, npv, 1, device)
        obs = _as_dev(obs, device).reshape(-1)
        if obs.numel() != self.npt:
            raise ValueError(f'obs has {obs.numel()} values, the data have {self.npt} times')
        sig = _as_dev(sigma, device).reshape(-1)
        if sig.numel() not in (1, npv):
            raise ValueError('sigma must be a scalar or one value per parameter vector')
        sig = sig.expand(npv).contiguous()
        tab, _k = self._table(device)
        d = self._device_data(device)
        lc = RRLcData(npt=self.npt, nlc=1, npb=1, nep=t0.shape[1], sorted=int(d['sorted']),
                      uniform_nsamples=d['uniform_ns'], times=ptr(d['times']), lcids=ptr(d['lcids']),
                      pbids=ptr(d['pbids']), epids=ptr(d['epids']), nsamples=ptr(d['nsamples']),
                      exptimes=ptr(d['exptimes']))
        prm = RRParams(npv=npv, kcols=1, t0cols=t0.shape[1], pad_=0, k=ptr(k), t0=ptr(t0), p=ptr(p), a=ptr(a),
                       inc=ptr(inc), e=ptr(e), w=ptr(w))
        lnl = torch.empty(npv, dtype=torch.float64, device=device)
        g = torch.empty((npv, 7), dtype=torch.float64, device=device) if grad else None
        byref = _lib.ctypes.byref
        check(lib.rr_lnlike(_PRECISIONS[self.precision], byref(tab), byref(lc), byref(prm), byref(ld), ptr(obs),
                            ptr(sig), ptr(lnl), ptr(g), stream_handle(device)))
        return lnl, g


RRModel = RoadRunnerModel
