"""Limb-darkening plug-in protocol.

Mirrors ``LDModel`` of ``pytransit/models/ldmodel.py:21-39``: a model is called as
``ldmodel(mu, x) -> (ldp, istar)`` with ``ldp[npv, npb, nmu]`` the profile at the
RoadRunner mu nodes and ``istar[npv, npb]`` its disk integral.  Subclasses define
``_evaluate(mu, x)`` and may override ``_integrate(x)``; the default integral is the
trapezoid ``2*pi*trapz(z*I(mu(z)), z)`` on ``niz`` uniform z nodes (ldmodel.py:33-39),
taken per (pv, passband).

Either returned array may be a numpy array or a torch tensor already on the GPU;
:class:`TabulatedLDModel` hands over device-resident profiles without a copy.
"""
from typing import Tuple

import numpy as np
from numpy import linspace, ndarray, pi, sqrt, trapezoid, zeros


class LDModel:
    def __init__(self, niz: int = 200):
        self._int_z = linspace(0, 1, niz)
        self._int_mu = sqrt(1 - self._int_z ** 2)

    def __call__(self, mu: ndarray, x: ndarray) -> Tuple[ndarray, ndarray]:
        return self._evaluate(mu, x), self._integrate(x)

    def _evaluate(self, mu: ndarray, x: ndarray) -> ndarray:
        raise NotImplementedError

    def _integrate(self, x: ndarray) -> ndarray:
        prof = np.asarray(self._evaluate(self._int_mu, x))
        prof = prof.reshape((-1,) + prof.shape[-2:]) if prof.ndim >= 2 else prof.reshape(1, 1, -1)
        npv, npb = prof.shape[0], prof.shape[1]
        ldi = zeros((npv, npb))
        for ipv in range(npv):
            for ipb in range(npb):
                ldi[ipv, ipb] = 2. * pi * trapezoid(self._int_z * prof[ipv, ipb], self._int_z)
        return ldi


class TabulatedLDModel(LDModel):
    """Profiles given directly as a table at the model's mu nodes.

    ``ldp`` is ``[npv, npb, nmu]`` and ``istar`` ``[npv, npb]`` (numpy or device
    tensors).  When ``istar`` is None it is computed on the GPU with the trapezoid
    of ``integrate_profiles_set`` (pytransit/models/numba/ldtkldm.py:77-91).
    The coefficient argument of ``__call__`` is ignored.
    """

    def __init__(self, ldp, istar=None):
        super().__init__()
        self.ldp = ldp
        self.istar = istar

    def __call__(self, mu, x):
        return self.ldp, self.istar


def _even_axis(nodes, name):
    """(x0, dx, n) of an evenly spaced axis given by its node values."""
    v = np.asarray(nodes, np.float64).ravel()
    if v.size < 2:
        raise ValueError(f'LDTkLDModel: the {name} axis needs at least 2 nodes')
    d = np.diff(v)
    if not (np.all(np.isfinite(v)) and np.all(d > 0) and np.allclose(d, d[0], rtol=1e-9, atol=0.0)):
        raise ValueError(f'LDTkLDModel: the {name} nodes must be finite, increasing and evenly spaced')
    return float(v[0]), float(v[1] - v[0]), int(v.size)


class LDTkLDModel(LDModel):
    """Limb darkening from a grid of stellar intensity profiles over (teff, logg,
    metallicity), interpolated trilinearly on the GPU at each parameter vector's
    stellar parameters; pytransit/models/ldtkldm.py:25-89.

    The reference builds the grid with LDTk (``LDPSetCreator`` over ``pbs`` and the
    ``teff``, ``logg`` and ``metal`` ranges) and resamples it to the model's mu
    nodes.  LDTk is not part of this package: as in the reference without LDTk,
    constructing from ``pbs``/``teff``/``logg``/``metal`` alone raises
    ``ModuleNotFoundError``.  Pass the grid instead: ``profiles`` is
    ``[nteff, nlogg, nmetal, npb, nmu]`` at the model's mu nodes, and ``teffs``,
    ``loggs`` and ``metals`` are the node values of the three axes (evenly
    spaced).  The grid is copied to each GPU once.

    Called as ``ldm(mu, x)`` with the stellar parameters ``x`` as ``[3]``,
    ``[npv, 3]`` or ``[npv, 1, 3]`` (teff, logg, metallicity; ldtkldm.py:78-83), it
    returns the profiles ``[npv, npb, nmu]`` (``rr_ldtk_interpolate``) and their disk
    integrals ``[npv, npb]`` (``rr_integrate_profiles``, the trapezoid of
    ldtkldm.py:77-91) as float64 tensors on the GPU.
    """

    def __init__(self, pbs=None, teff=None, logg=None, metal=None, cache=None, dataset: str = 'vis-lowres', *,
                 profiles=None, teffs=None, loggs=None, metals=None):
        super().__init__()
        if profiles is None:
            raise ModuleNotFoundError('LDTkLDModel requires LDTk; without it, give the profile grid as '
                                      'profiles=[nteff, nlogg, nmetal, npb, nmu] with its teffs, loggs and metals nodes.')
        self.pbs = pbs
        self.profiles = profiles
        shape = tuple(profiles.shape)
        if len(shape) != 5:
            raise ValueError(f'LDTkLDModel: profiles must be [nteff, nlogg, nmetal, npb, nmu], got shape {shape}')
        self.teff0, self.dteff, self.nteff = _even_axis(teffs, 'teff')
        self.logg0, self.dlogg, self.nlogg = _even_axis(loggs, 'logg')
        self.metal0, self.dmetal, self.nmetal = _even_axis(metals, 'metallicity')
        if shape[:3] != (self.nteff, self.nlogg, self.nmetal):
            raise ValueError(f'LDTkLDModel: profiles have grid shape {shape[:3]}, the axes give '
                             f'{(self.nteff, self.nlogg, self.nmetal)}')
        self.npb, self.nmu = shape[3], shape[4]
        self._grid = {}

    def _device_grid(self, device):
        import torch
        g = self._grid.get(device)
        if g is None:
            g = torch.as_tensor(self.profiles, dtype=torch.float64, device=device).contiguous()
            self._grid[device] = g
        return g

    def __call__(self, mu, x):
        import torch
        from . import _lib
        mu = np.asarray(mu, np.float64)
        if mu.size != self.nmu:
            raise ValueError(f'LDTkLDModel: the profiles have {self.nmu} mu nodes, the model has {mu.size}')
        lib = _lib.load()
        if isinstance(x, torch.Tensor) and x.is_cuda:
            device = x.device
        else:
            device = torch.device('cuda', torch.cuda.current_device())
        par = torch.as_tensor(x, dtype=torch.float64, device=device)
        if par.dim() == 1:
            par = par.reshape(1, 1, -1)
        elif par.dim() == 2:
            par = par.reshape(par.shape[0], 1, par.shape[1])
        if par.dim() != 3 or par.shape[2] < 3:
            raise ValueError(f'LDTkLDModel: the stellar parameters must be [npv, 3] (teff, logg, metallicity), '
                             f'got shape {tuple(par.shape)}')
        par = par[:, 0, :3].contiguous()
        npv = par.shape[0]
        grid = self._device_grid(device)
        ldp = torch.empty((npv, self.npb, self.nmu), dtype=torch.float64, device=device)
        istar = torch.empty((npv, self.npb), dtype=torch.float64, device=device)
        mu_d = torch.as_tensor(mu, device=device)
        stream = _lib.stream_handle(device)
        _lib.check(lib.rr_ldtk_interpolate(_lib.ptr(grid), self.nteff, self.nlogg, self.nmetal, self.npb, self.nmu,
                                           self.teff0, self.dteff, self.logg0, self.dlogg, self.metal0, self.dmetal,
                                           _lib.ptr(par), 3, npv, _lib.ptr(ldp), stream))
        _lib.check(lib.rr_integrate_profiles(_lib.ptr(mu_d), self.nmu, _lib.ptr(ldp), npv, self.npb,
                                             _lib.ptr(istar), stream))
        return ldp, istar
