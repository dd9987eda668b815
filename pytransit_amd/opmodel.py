"""OblatePlanetModel on MI355X: the RoadRunner model for a planet whose sky
projection is an ellipse.

Drop-in for ``pytransit.models.roadrunner.opmodel.OblatePlanetModel``
(opmodel.py:43-111, kernel opmodel_full.py:40-134): the RoadRunnerModel
constructor plus ``nlines`` (the number of scanlines of the ellipse), and
``evaluate(k, f, alpha, ldc, t0, p, a, i, e=0.0, w=0.0, copy=True)`` with the
flattening ``f`` and the sky-plane rotation ``alpha`` per pv.

The per-pv setup (validity, Taylor orbit, contact window, limb-darkening means
and disk integrals) is the RoadRunner setup kernel's; the flux kernel
(``op_flux_kernel``, FP64) averages (I* - I(g) A_ellipse) / I* over the exposure,
with the area of the ellipse inside the stellar disk summed over its scanlines
(ecintersection.py:137-207).  The window margin is 0.0015 d + the exposure time
(opmodel_full.py:101-104).
"""
from typing import Callable, List, Tuple, Union

import torch
from numpy import ndarray

from . import _lib
from ._lib import check, ptr, stream_handle
from .rrmodel import RoadRunnerModel, _as_dev

__all__ = ['OblatePlanetModel']


class OblatePlanetModel(RoadRunnerModel):
    def __init__(self, ldmodel: Union[str, Callable, Tuple[Callable, Callable]] = 'quadratic',
                 precompute_weights: bool = False, klims: tuple = (0.005, 0.5), nk: int = 256,
                 nzin: int = 20, nzlimb: int = 20, zcut: float = 0.7, ng: int = 100, nlines: int = 100,
                 nthreads: int = 1, small_planet_limit: float = 0.05, device=None, **kwargs):
        """opmodel.py:44-48.  The model is evaluated in FP64 (the reference's arithmetic)."""
        super().__init__(ldmodel, precompute_weights, klims, nk, nzin, nzlimb, zcut, ng, nthreads,
                         small_planet_limit, precision='fp64', device=device, **kwargs)
        self.nlines = nlines

    def evaluate(self, k: Union[float, ndarray], f: Union[float, ndarray], alpha: Union[float, ndarray],
                 ldc: Union[ndarray, List], t0: Union[float, ndarray], p: Union[float, ndarray],
                 a: Union[float, ndarray], i: Union[float, ndarray], e: Union[float, ndarray] = 0.0,
                 w: Union[float, ndarray] = 0.0, copy: bool = True):
        """Transit of an oblate planet (opmodel.py:50-111): squeezed flux ``[npv, npt]``.

        ``f`` is the flattening (semi-minor axis k (1 - f)) and ``alpha`` the rotation
        of the ellipse's major axis on the sky, one value each per pv.  Invalid
        parameter vectors give NaN rows.
        """
        flux = self._evaluate_op(k, f, alpha, ldc, t0, p, a, i, e, w).squeeze()
        return _lib.to_host(flux) if copy else flux

    def __call__(self, *args, **kwargs):
        return self.evaluate(*args, **kwargs)

    def _evaluate_op(self, k, f, alpha, ldc, t0, p, a, i, e=0.0, w=0.0):
        lib = _lib.load()
        dev = self.device
        npv, tab, lc, prm, ld, _keep = self._descriptors(k, ldc, t0, p, a, i, e, w)
        fl = _as_dev(f, dev).reshape(-1)
        al = _as_dev(alpha, dev).reshape(-1)
        for name, x in (('f', fl), ('alpha', al)):
            if x.numel() not in (1, npv):
                raise ValueError(f"{name} has {x.numel()} values, expected 1 or npv={npv}")
        fl = fl.expand(npv).contiguous()
        al = al.expand(npv).contiguous()
        nbytes = lib.rr_workspace_bytes(npv, self.npt, self.npb, self.ng, self.nlc)
        ws = self._workspace(dev, nbytes)
        flux = torch.empty((npv, self.npt), dtype=torch.float64, device=dev)
        byref = _lib.ctypes.byref
        check(lib.op_evaluate(byref(tab), byref(lc), byref(prm), byref(ld), ptr(fl), ptr(al), int(self.nlines),
                              ptr(ws), ws.numel(), ptr(flux), stream_handle(dev)))
        return flux
