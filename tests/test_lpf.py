"""The LPF-layout wrapper (pytransit/lpf/lpf.py:85-91,435-443): map_ldc on the CPU,
and on the GPU the wrapper's flux and likelihood against RoadRunnerModel's."""
import math

import numpy as np
import pytest
import torch

from pytransit_amd import RoadRunnerModel, workloads
from pytransit_amd.lpf import RoadRunnerLPFModel, _pairs, map_ldc


def test_map_ldc_kipping_parametrisation():
    # q1 = 0.36, q2 = 0.25: sqrt(q1) = 0.6, u = 2 * 0.6 * 0.25 = 0.3, v = 0.6 * 0.5 = 0.3
    uv = map_ldc([[0.36, 0.25, 0.64, 0.5]])
    np.testing.assert_allclose(uv, [[0.3, 0.3, 0.8, 0.0]], rtol=1e-15)
    # q1 = 1 covers the full triangle edge: u + v = 1
    u, v = map_ldc([1.0, 0.3])[0]
    assert u + v == pytest.approx(1.0, abs=1e-15)


def test_pairs_layout():
    c = np.arange(12.0).reshape(2, 6)
    p = _pairs(c, 3)
    assert p.shape == (2, 3, 2) and p[1, 2].tolist() == [10.0, 11.0]
    with pytest.raises(ValueError, match='two limb darkening'):
        _pairs(c, 2)


@pytest.mark.gpu
def test_lpf_layout_flux_and_likelihood(gpu):
    print('case: LPF layouts (ldc [npv, 2 npb], 1-D t0) vs RoadRunnerModel', flush=True)
    wl = workloads.rr_small(npv=12, npt=900)
    npb = 2
    lcids = (np.arange(wl.times.size) * npb) // wl.times.size
    rng = np.random.default_rng(4)
    ldc3 = rng.uniform(0.1, 0.4, (12, npb, 2))
    ref = RoadRunnerModel('quadratic')
    ref.set_data(wl.times, lcids=lcids, pbids=np.arange(npb))
    f_ref = ref.evaluate(wl.k, ldc3, wl.t0, wl.p, wl.a, wl.i)
    tm = RoadRunnerLPFModel()
    tm.set_data(wl.times, lcids=lcids, pbids=np.arange(npb))
    f = tm.evaluate(wl.k, ldc3.reshape(12, 2 * npb), wl.t0[:, 0], wl.p, wl.a, wl.i)
    np.testing.assert_array_equal(f, f_ref)
    # fused white-noise likelihood on one light curve
    one = RoadRunnerLPFModel()
    one.set_data(wl.times)
    obs = one.evaluate(wl.k[:1], ldc3[:1, 0], wl.t0[:1, 0], wl.p[:1], wl.a[:1], wl.i[:1])
    obs = np.asarray(obs, float) + rng.normal(0, 1e-3, wl.times.size)
    lnl = one.lnlikelihood(wl.k, ldc3[:, 0], wl.t0[:, 0], wl.p, wl.a, wl.i, obs, 1e-3)
    fm = np.asarray(one.evaluate(wl.k, ldc3[:, 0], wl.t0[:, 0], wl.p, wl.a, wl.i), float)
    r = (obs[None, :] - fm) / 1e-3
    expect = (-math.log(1e-3) - 0.5 * math.log(2 * math.pi) - 0.5 * r ** 2).sum(1)
    np.testing.assert_allclose(lnl, expect, rtol=1e-10)
