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
