"""The fused log-likelihood and its hand-written adjoint against the FP64 oracle.

lnL is checked against lnlike_normal's form (wnloglikelihood.py:22-35) applied to
the oracle's flux; the gradient d lnL / d(k, t0, p, a, i, u1, u2) against central
finite differences (h = 1e-6) of that oracle lnL.  Bar: |g - g_fd| <= 1e-4 (|g_fd|
+ s_j), where s_j is the RMS of the FD gradient of parameter j over the population
(so a component that happens to be near zero is judged on its parameter's scale).
"""
import math

import numpy as np
import pytest
import torch

from pytransit_amd import RoadRunnerModel, workloads
from pytransit_amd.autograd import rr_lnlike

pytestmark = pytest.mark.gpu

NAMES = ('k', 't0', 'p', 'a', 'i', 'u1', 'u2')


def say(*a):
    print(*a, flush=True)


def setup_case(oracle, npv):
    wl = workloads.rr_gradient(npv=npv)
    tr = wl.truth
    om = oracle.OracleRR('quadratic')
    om.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    f_true = om.evaluate(np.array([[tr['k']]]), np.array([[[tr['u1'], tr['u2']]]]), np.array([[tr['t0']]]),
                         tr['p'], tr['a'], tr['i'])[0]
    obs = f_true + wl.noise_draw
    theta = np.stack([wl.pop[n] for n in NAMES], 1)
    return wl, om, obs, theta


def oracle_lnl(om, obs, sigma, th):
    f = om.evaluate(th[:, 0:1], th[:, 5:7].reshape(-1, 1, 2), th[:, 1:2], th[:, 2], th[:, 3], th[:, 4])
    r = (obs[None, :] - f) / sigma
    return (-math.log(sigma) - 0.5 * math.log(2 * math.pi) - 0.5 * r ** 2).sum(1)


def fd_grad(om, obs, sigma, theta, h=1e-6):
    g = np.zeros_like(theta)
    for j in range(theta.shape[1]):
        tp, tm = theta.copy(), theta.copy()
        tp[:, j] += h
        tm[:, j] -= h
        g[:, j] = (oracle_lnl(om, obs, sigma, tp) - oracle_lnl(om, obs, sigma, tm)) / (2 * h)
    return g


def gpu_lnl(wl, prec, obs, sigma, theta, grad=True):
    tm = RoadRunnerModel('quadratic', precision=prec)
    tm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    return tm.lnlike(theta[:, 0], theta[:, 5:7].reshape(-1, 1, 2), theta[:, 1], theta[:, 2], theta[:, 3],
                     theta[:, 4], obs, sigma, grad=grad)


@pytest.fixture(scope='module')
def case(oracle):
    wl, om, obs, theta = setup_case(oracle, npv=10)
    sigma = wl.noise
    g_fd = {h: fd_grad(om, obs, sigma, theta, h) for h in (1e-6, 1e-8)}
    lnl_o = oracle_lnl(om, obs, sigma, theta)
    return wl, om, obs, theta, sigma, g_fd, lnl_o


# The model is piecewise linear in g (limb-darkening lerp) and k (table lerp): a
# central difference that straddles a knot measures a secant, the adjoint the
# one-sided slope.  At h = 1e-6 ~1e-3 of the samples straddle a g knot (~2e-5
# relative on t0 and i); h = 1e-8 straddles ~100x fewer and pins the algebra.
@pytest.mark.parametrize('prec,h,lnl_rtol,g_tol', [('fp64', 1e-6, 1e-12, 1e-4), ('mixed', 1e-6, 1e-6, 1e-4),
                                                  ('fp64', 1e-8, 1e-12, 5e-6)])
def test_lnlike_and_gradient_vs_finite_differences(gpu, case, prec, h, lnl_rtol, g_tol):
    wl, om, obs, theta, sigma, g_fds, lnl_o = case
    g_fd = g_fds[h]
    say(f'case: lnL + adjoint, 10 pv x {wl.times.size} t, nss=5, {prec}, FD h={h:g}')
    lnl, g = gpu_lnl(wl, prec, obs, sigma, theta)
    np.testing.assert_allclose(lnl, lnl_o, rtol=lnl_rtol)
    scale = np.sqrt((g_fd ** 2).mean(0))
    rel = np.abs(g - g_fd) / (np.abs(g_fd) + scale)
    for j, n in enumerate(NAMES):
        say(f'  {n:3s} max rel err {rel[:, j].max():.2e}  (|g| ~ {scale[j]:.3e})')
    assert (scale > 0).all()
    assert rel.max() <= g_tol


def test_autograd_backward_is_the_adjoint(gpu, case):
    wl, om, obs, theta, sigma, _, lnl_o = case
    say('case: torch.autograd through RRLogLikelihood')
    tm = RoadRunnerModel('quadratic', precision='fp64')
    tm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    th = torch.tensor(theta, device=gpu, requires_grad=True)
    obs_d = torch.as_tensor(obs, device=gpu)
    lnl = rr_lnlike(th, tm, obs_d, sigma)
    w = torch.linspace(0.5, 1.5, lnl.numel(), device=gpu, dtype=lnl.dtype)
    (lnl * w).sum().backward()
    _, g = gpu_lnl(wl, 'fp64', obs, sigma, theta)
    np.testing.assert_allclose(th.grad.cpu().numpy(), g * w.cpu().numpy()[:, None], rtol=1e-12)


def test_lnlike_invalid_and_out_of_table(gpu, case):
    wl, om, obs, theta, sigma, _, lnl_o = case
    say('case: invalid pv (a <= 1) and k outside klims give NaN')
    th = theta[:3].copy()
    th[1, 3] = 0.8
    th[2, 0] = 0.7
    lnl, g = gpu_lnl(wl, 'fp64', obs, sigma, th)
    assert np.isfinite(lnl[0]) and np.isnan(lnl[1]) and np.isnan(lnl[2])
    assert np.isnan(g[1]).all() and np.isnan(g[2]).all()


# At the gradient benchmark's size (2048 pv x 10 000 t, nss = 5).  Two plain
# relative checks on the components with |g| >= 0.3 x the parameter's RMS over
# the population (about 3/4 of them; below that floor a component is a near-zero
# difference of large per-point terms and its relative error says nothing):
#   * the mixed-precision adjoint against the FP64 adjoint, bar 1e-4;
#   * the FP64 adjoint against central differences of the FP64 lnL on the GPU
#     (pinned to the oracle's lnL on a 16-pv slice here, and at rtol 1e-12 above)
#     at h = 1e-8 on 256 pv.  At h = 1e-6 the secants straddle the limb-darkening
#     lerp's knots often enough at this size to move t0 and i by up to 7e-4.
FLOOR = 0.3


@pytest.fixture(scope='module')
def scale_case(gpu):
    wl = workloads.rr_gradient(npv=2048)
    tr = wl.truth
    ref = RoadRunnerModel('quadratic', precision='fp64')
    ref.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    f_true = ref.evaluate([[tr['k']]], [[[tr['u1'], tr['u2']]]], [[tr['t0']]], tr['p'], tr['a'], tr['i'])
    obs = f_true + wl.noise_draw
    theta = np.stack([wl.pop[n] for n in NAMES], 1)
    return wl, obs, theta


def plain_rel_above_floor(g, g_ref):
    scale = np.sqrt((g_ref ** 2).mean(0))
    keep = np.abs(g_ref) >= FLOOR * scale
    rel = np.where(keep, np.abs(g - g_ref) / np.where(keep, np.abs(g_ref), 1.0), 0.0)
    return rel.max(0), keep.mean(0)


def test_mixed_adjoint_vs_fp64_at_scale(gpu, scale_case):
    wl, obs, theta = scale_case
    say(f'case: mixed vs FP64 adjoint, {theta.shape[0]} pv x {wl.times.size} t, nss=5')
    lnl_m, g_m = gpu_lnl(wl, 'mixed', obs, wl.noise, theta)
    lnl_d, g_d = gpu_lnl(wl, 'fp64', obs, wl.noise, theta)
    assert np.isfinite(g_d).all() and np.isfinite(g_m).all()
    rel, kept = plain_rel_above_floor(g_m, g_d)
    for j, n in enumerate(NAMES):
        say(f'  {n:3s} max plain rel {rel[j]:.2e} over {kept[j]:.0%} of the pv')
    assert (kept > 0.5).all()
    assert rel.max() <= 1e-4
    np.testing.assert_allclose(lnl_m, lnl_d, rtol=1e-6)


def test_fp64_adjoint_vs_finite_differences_at_scale(gpu, scale_case, oracle):
    wl, obs, theta = scale_case
    th = theta[:256]
    say(f'case: FP64 adjoint vs central differences (h = 1e-8), 256 pv x {wl.times.size} t, nss=5')
    lnl, g = gpu_lnl(wl, 'fp64', obs, wl.noise, th)
    # the differentiated function is the oracle's lnL
    om = oracle.OracleRR('quadratic')
    om.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    np.testing.assert_allclose(lnl[:16], oracle_lnl(om, obs, wl.noise, th[:16]), rtol=1e-12)
    h = 1e-8
    g_fd = np.zeros_like(th)
    for j in range(th.shape[1]):
        tp, tm = th.copy(), th.copy()
        tp[:, j] += h
        tm[:, j] -= h
        g_fd[:, j] = (gpu_lnl(wl, 'fp64', obs, wl.noise, tp, grad=False) -
                      gpu_lnl(wl, 'fp64', obs, wl.noise, tm, grad=False)) / (2 * h)
    rel, kept = plain_rel_above_floor(g, g_fd)
    for j, n in enumerate(NAMES):
        say(f'  {n:3s} max plain rel {rel[j]:.2e} over {kept[j]:.0%} of the pv')
    assert (kept > 0.5).all()
    assert rel.max() <= 1e-4
