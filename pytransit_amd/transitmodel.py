"""Light-curve data handling shared by the RoadRunner models.

Mirrors ``TransitModel`` of ``pytransit/models/transitmodel.py:36-135``: the same
``set_data`` arguments, defaults, validation and error messages.  The validated
arrays are kept on the host and uploaded once per device (times as float64,
indices as int32) the first time a model is evaluated there; a later ``set_data``
replaces them.
"""
from typing import List, Optional, Union

import numpy as np
import torch
from numpy import asarray, atleast_1d, float64, integer, issubdtype, ndarray, ones, unique, zeros


def runs_sorted(time: ndarray, lcids: ndarray) -> bool:
    """True if the light curves are contiguous (lcids non-decreasing) and each is
    sorted in time: the layout the flux kernel classifies run by run."""
    if time.size < 2:
        return True
    dl = np.diff(lcids)
    return bool(np.all(dl >= 0) and np.all(np.diff(time)[dl == 0] >= 0))


def layout_flags(time: ndarray, lcids: ndarray, epids: ndarray, exptimes: ndarray) -> int:
    """The rr_lcdata.sorted bits: 1 (RR_SORTED_RUNS) when runs_sorted; with it 2
    (RR_SORTED_TIMES) when the times are sorted over the whole array and 4
    (RR_ONE_WINDOW) when every light curve has the same epoch index and exposure
    time, i.e. a pv has one transit window for all light curves."""
    if not runs_sorted(time, lcids):
        return 0
    flags = 1
    if time.size < 2 or bool(np.all(np.diff(time) >= 0)):
        flags |= 2
    if (epids.size == 0 or (epids == epids[0]).all()) and (exptimes.size == 0 or (exptimes == exptimes[0]).all()):
        flags |= 4
    return flags


class TransitModel:
    """Exoplanet transit light curve model (transitmodel.py:36-54)."""

    def __init__(self) -> None:
        self.time_id: Optional[int] = None
        self.time: Optional[ndarray] = None
        self.lcids: Optional[ndarray] = None
        self.pbids: Optional[ndarray] = None
        self.nsamples: Optional[ndarray] = None
        self.exptimes: Optional[ndarray] = None
        self.epids: Optional[ndarray] = None
        self.nlc: int = 0
        self.npt: int = 0
        self.npb: int = 0
        self._dev_data = {}

    def set_data(self, time: Union[ndarray, List],
                 lcids: Optional[Union[ndarray, List]] = None,
                 pbids: Optional[Union[ndarray, List]] = None,
                 nsamples: Optional[Union[ndarray, List]] = None,
                 exptimes: Optional[Union[ndarray, List]] = None,
                 epids: Optional[Union[ndarray, List]] = None) -> None:
        """Set the data for the transit model (transitmodel.py:56-125).

        Parameters
        ----------
        time : array-like
            Mid-exposure times at which the model is evaluated.
        lcids : array-like, optional
            Integer light-curve index of each time; same size as ``time``.
        pbids : array-like, optional
            Passband index of each light curve; ``pbids.size == unique(lcids).size``.
        nsamples : int or array-like, optional
            Supersamples per exposure, for all light curves or one per light curve.
        exptimes : float or array-like, optional
            Exposure times, for all light curves or one per light curve.
        epids : array-like, optional
            Epoch (zero-epoch/period set) index of each light curve.
        """
        if (id(time) == self.time_id and lcids is None and pbids is None and nsamples is None
                and exptimes is None and epids is None):
            return

        if isinstance(time, torch.Tensor):
            time = time.detach().cpu().numpy()
        self.time_id = id(time)
        self.time = asarray(time, float64)
        self.npt = self.time.size

        self.lcids = asarray(lcids) if lcids is not None else zeros(self.npt, 'int')
        self.nlc = unique(self.lcids).size
        if not issubdtype(self.lcids.dtype, integer):
            raise ValueError(f"The light curve indices must be given as integers instead of {self.lcids.dtype}.")
        if self.lcids.size != self.npt:
            raise ValueError(f"Light curve index array size ({self.lcids.size}) should equal to the number of "
                             f"datapoints ({self.npt}).")

        self.pbids = asarray(pbids) if pbids is not None else zeros(self.nlc, 'int')
        self.npb = unique(self.pbids).size
        if not issubdtype(self.pbids.dtype, integer):
            raise ValueError(f"The passband indices must be given as integers instead of {self.pbids.dtype}.")
        if self.pbids.size != self.nlc:
            raise ValueError(f"Passband index array size ({self.pbids.size}) should equal to the number of ligt "
                             f"curves ({self.nlc}).")
        if not (self.pbids.max() == (self.npb - 1) and self.pbids.min() == 0):
            raise ValueError(f"Passband indices (`pbids`) for {self.npb} unique passbands should be given as "
                             f"integers between 0 and {self.npb - 1}.")

        self.epids = asarray(epids) if epids is not None else zeros(self.nlc, 'int')
        self.nsamples = atleast_1d(nsamples) if nsamples is not None else ones(self.nlc, 'int')
        self.exptimes = atleast_1d(exptimes) if exptimes is not None else zeros(self.nlc, 'int')
        self._validate_device_bounds()
        self._dev_data = {}

    def _validate_device_bounds(self):
        """Index ranges the GPU kernels rely on (the reference does not check these)."""
        if self.npt and (self.lcids.min() < 0 or self.lcids.max() >= self.nlc):
            raise ValueError(f"Light curve indices should be integers between 0 and {self.nlc - 1}.")
        if self.epids.size and self.epids.min() < 0:
            raise ValueError("Epoch indices must be non-negative.")
        if self.npt > 2 ** 31 - 1:
            raise ValueError("At most 2**31-1 data points are supported.")

    def _per_lc(self, x, dtype):
        """Broadcast a scalar or per-light-curve array to [nlc] (model_full.py:21-24)."""
        out = np.zeros(self.nlc, dtype)
        out[:] = x
        return out

    def _device_data(self, device: torch.device):
        """Device copies of the light-curve arrays (uploaded once per set_data)."""
        key = str(device)
        d = self._dev_data.get(key)
        if d is None:
            if self.time is None:
                raise ValueError('set_data() must be called before evaluate()')
            nsamples = self._per_lc(self.nsamples, np.int64)
            if (nsamples < 1).any():
                raise ValueError('nsamples must be >= 1')
            d = dict(
                sorted=layout_flags(self.time, self.lcids, self.epids, self._per_lc(self.exptimes, np.float64)),
                uniform_ns=int(nsamples[0]) if (nsamples == nsamples[0]).all() else 0,
                times=torch.as_tensor(np.ascontiguousarray(self.time), device=device),
                lcids=torch.as_tensor(self.lcids.astype(np.int32), device=device),
                pbids=torch.as_tensor(self.pbids.astype(np.int32), device=device),
                epids=torch.as_tensor(self.epids.astype(np.int32), device=device),
                nsamples=torch.as_tensor(nsamples.astype(np.int32), device=device),
                exptimes=torch.as_tensor(self._per_lc(self.exptimes, np.float64), device=device),
            )
            self._dev_data[key] = d
        return d

    def __call__(self, *nargs, **kwargs):
        raise NotImplementedError

    def evaluate(self, k, ldc, t0, p, a, i, e=None, w=None, copy: bool = True):
        raise NotImplementedError
