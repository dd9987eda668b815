"""GPU parity of the eclipse, eclipse-spectroscopy and oblate-planet models
(ec_kernels.hip) against the FP64 oracle (oracle/ec_oracle.py).  Bar: 1e-9
absolute (FP64, as the reference)."""
import math

import numpy as np
import pytest

from pytransit_amd import EclipseModel, EclipseSpectroscopyModel, OblatePlanetModel

pytestmark = pytest.mark.gpu

TOL = 1e-9


def say(*a):
    print(*a, flush=True)


def test_eclipse_model_reference_test_case(gpu):
    # tests/test_roadrunner_eclipse.py:16-69 through the class API
    from oracle import ec_oracle as eo
    say('case: EclipseModel, the reference test geometry (k=0.1, a=8, p=2, i=90 deg), 5000 t')
    times = np.linspace(-0.5, 3.5, 5000)
    em = EclipseModel()
    em.set_data(times)
    f = em.evaluate(0.1, 0.0, 2.0, 8.0, 0.5 * math.pi)
    assert f.shape == (times.size,)
    base = math.pi * 0.01
    assert f.max() == pytest.approx(base, abs=1e-6)
    assert f[np.argmin(np.abs(times - 1.0))] < base - 1e-4 and f.min() < 1e-3
    ref = eo.eclipse_model(times, np.array([0.1]), np.zeros((1, 1)), np.array([2.0]), np.array([8.0]),
                           np.array([0.5 * math.pi]), np.zeros(1), np.zeros(1), 1.0, 1,
                           np.zeros(times.size, np.int64), np.zeros(1, np.int64), np.ones(1, np.int64),
                           np.zeros(1))[0]
    err = np.abs(f - ref).max()
    say(f'  max |dF| = {err:.2e}')
    assert err <= TOL


def test_eclipse_model_population_against_oracle(gpu):
    from oracle import ec_oracle as eo
    say('case: EclipseModel, 6 pv (eccentric, one invalid), 2 light curves, nss=3 / 1')
    rng = np.random.default_rng(7)
    npv = 6
    times = np.concatenate([np.linspace(0.7, 1.4, 700), np.linspace(2.6, 3.5, 500)])
    lcids = np.r_[np.zeros(700, int), np.ones(500, int)]
    k = rng.uniform(0.06, 0.14, npv)
    p = rng.uniform(1.8, 2.4, npv)
    a = rng.uniform(5.0, 10.0, npv)
    inc = np.arccos(rng.uniform(0.0, 0.7, npv) / a)
    e = np.array([0.0, 0.1, 0.25, 0.0, 0.05, 0.3])
    w = rng.uniform(-math.pi, math.pi, npv)
    a[4] = 0.9                       # invalid: a NaN row
    t0 = np.stack([rng.normal(0, 0.01, npv), rng.normal(0, 0.01, npv)], 1)
    em = EclipseModel()
    em.set_data(times, lcids=lcids, pbids=[0, 1], nsamples=[3, 1], exptimes=[0.01, 0.0])
    got = em.evaluate(k, t0, p, a, inc, e, w, rstar=1.3)
    # the reference passes pbids as the kernel's epoch indices (new_eclipse_model.py:70)
    ref = eo.eclipse_model(times, k, t0, p, a, inc, e, w, 1.3, 2, lcids, np.array([0, 1]), np.array([3, 1]),
                           np.array([0.01, 0.0]))
    assert np.isnan(got[4]).all() and np.isnan(ref[4]).all()
    m = np.arange(npv) != 4
    err = np.abs(got[m] - ref[m]).max()
    say(f'  max |dF| = {err:.2e}, eclipse points {(ref[m] < math.pi * k[m, None] ** 2 - 1e-12).sum()}')
    assert (ref[m] < math.pi * k[m, None] ** 2 - 1e-12).sum() > 500
    assert err <= TOL


def test_eclipse_spectroscopy_against_oracle(gpu):
    from oracle import ec_oracle as eo
    say('case: EclipseSpectroscopyModel, 4 pv x 7 pb x 1500 t, nss=2')
    rng = np.random.default_rng(8)
    npv, npb = 4, 7
    times = np.linspace(1.3, 1.9, 1500)
    k = rng.uniform(0.07, 0.13, npv)
    p = np.full(npv, 3.2)
    a = rng.uniform(7.0, 11.0, npv)
    inc = np.arccos(rng.uniform(0.0, 0.6, npv) / a)
    e = np.array([0.0, 0.15, 0.0, 0.3])
    w = np.array([0.0, 1.0, 0.0, -0.5])
    rs = rng.uniform(0.7, 1.4, npv)
    fr = rng.uniform(1e-4, 5e-3, (npv, npb))
    t0 = rng.normal(0.0, 0.01, npv)
    es = EclipseSpectroscopyModel()
    es.set_data(times, nsamples=2, exptimes=0.01)
    got = es.evaluate(fr, k, t0, p, a, inc, e, w, rs)
    ref = eo.esmodel(times, k, t0, p, a, inc, e, w, rs, fr, 2, 0.01)
    assert got.shape == ref.shape == (npv, npb, times.size)
    err = np.abs(got - ref).max()
    say(f'  max |dF| = {err:.2e}, in-eclipse points {(ref < 1).sum()}')
    assert (ref < 1).sum() > 5000
    assert err <= TOL


@pytest.mark.parametrize('nlines', [100, 37])
def test_oblate_planet_against_oracle(gpu, oracle, nlines):
    from oracle import ec_oracle as eo
    say(f'case: OblatePlanetModel, 3 pv x 301 t, quadratic, nss=2, {nlines} scanlines, one off-table k')
    npv = 3
    times = np.linspace(-0.12, 0.12, 301)
    k = np.array([[0.09], [0.12], [0.6]])        # the last k is beyond klims: exact 2-D weights
    f = np.array([0.1, 0.3, 0.05])
    alpha = np.array([0.3, -1.1, 2.0])
    ldc = np.array([[[0.4, 0.2]], [[0.3, 0.1]], [[0.5, 0.25]]])
    t0 = np.array([[0.001], [-0.002], [0.0]])
    p, a, inc = np.array([2.0, 3.0, 4.0]), np.array([7.0, 9.0, 11.0]), np.array([1.55, 1.53, 1.565])
    e, w = np.array([0.0, 0.1, 0.0]), np.array([0.0, 0.4, 0.0])
    om = OblatePlanetModel('quadratic', nlines=nlines)
    om.set_data(times, nsamples=2, exptimes=0.005)
    got = om.evaluate(k, f, alpha, ldc, t0, p, a, inc, e, w)
    orr = oracle.OracleRR('quadratic')
    ldp, istar = orr.ld(ldc)
    ref = eo.opmodel(times, k, f, alpha, t0, p, a, inc, e, w, 1, 1, 1, nlines, np.zeros(times.size, np.int64),
                     np.zeros(1, np.int64), np.zeros(1, np.int64), np.array([2]), np.array([0.005]), ldp, istar,
                     orr.weights, orr.dk, 0.005, 0.5, orr.dg, orr.ze)
    err = np.abs(got - ref).max()
    say(f'  max |dF| = {err:.2e}, in-transit points {(ref < 1).sum()}')
    assert (ref < 1).sum() > 300
    assert err <= TOL
