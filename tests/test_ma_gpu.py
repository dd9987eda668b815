"""QuadraticModel (Mandel & Agol) on the GPU against the FP64 oracle's
quadratic_model (ma_quadratic_nb.py:532-595 restated; its eval_quad_z is pinned by
the test_ma_quadratic_nb.py values in test_oracle.py).  Bar: 1e-9 absolute."""
import numpy as np
import pytest

from pytransit_amd import QuadraticModel, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_ma_population_multi_passband(gpu, oracle):
    say('case: MA 20 pv x 1500 t, 2 light curves / passbands, nss 1 and 4, NaN k and ldc')
    wl = workloads.rr_small(npv=20, npt=1500, seed=7)
    lcids = (np.arange(wl.times.size) * 2) // wl.times.size
    pbids, epids = np.array([0, 1]), np.array([0, 0])
    nsamples, exptimes = np.array([1, 4]), np.array([0.0, 0.01])
    rng = np.random.default_rng(7)
    k = np.concatenate([wl.k, wl.k * 1.02], 1)
    ldc = rng.uniform(0.1, 0.4, (20, 4))
    k[3, 1] = np.nan          # inf row
    ldc[5, 2] = np.nan        # inf points of passband 1
    fo = oracle.quadratic_model(wl.times, k, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w, ldc, lcids, pbids, epids,
                                nsamples, exptimes)
    tm = QuadraticModel()
    tm.set_data(wl.times, lcids=lcids, pbids=pbids, nsamples=nsamples, exptimes=exptimes, epids=epids)
    fg = tm.evaluate(k, ldc, wl.t0[:, 0], wl.p, wl.a, wl.i)
    assert fg.shape == fo.shape
    assert np.array_equal(np.isinf(fg), np.isinf(fo)) and np.isinf(fg[3]).all()
    fin = np.isfinite(fo)
    err = np.abs(fg[fin] - fo[fin]).max()
    say(f'  max |dF| = {err:.3e}, in-transit points {(fo[fin] < 1).sum()}')
    assert (fo[fin] < 1).sum() > 2000
    assert err <= 1e-9


def test_ma_scalar_parameters(gpu, oracle):
    say('case: MA scalar parameters')
    t = np.linspace(-0.2, 0.2, 801)
    tm = QuadraticModel()
    tm.set_data(t)
    fg = tm.evaluate(0.12, [0.35, 0.2], 0.001, 3.2, 9.5, 1.53)
    fo = oracle.quadratic_model(t, [[0.12]], [[0.001]], [3.2], [9.5], [1.53], [0.0], [0.0], [[0.35, 0.2]],
                                np.zeros(801, int), [0], [0], [1], [0.0])[0]
    assert fg.shape == (801,)
    assert np.abs(fg - fo).max() <= 1e-9
    assert fg.min() < 0.99
