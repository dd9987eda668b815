"""GPU parity of the RoadRunner path against the FP64 oracle.

Bars: FP64 path <= 1e-9 absolute flux (identical NaN pattern); mixed path
<= 1 ppm absolute flux.  Every test prints its case before touching the GPU.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from pytransit_amd import RoadRunnerModel, TabulatedLDModel, _lib, workloads

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_TOL = 1e-6


def say(*a):
    print(*a, flush=True)


def oracle_flux(oracle, wl, ldname, npv=None, nzin=20, nzlimb=20, lcids=None, pbids=None, epids=None):
    om = oracle.OracleRR(ldname, nzin=nzin, nzlimb=nzlimb)
    om.set_data(wl.times, lcids=lcids, pbids=pbids, nsamples=wl.nsamples, exptimes=wl.exptimes, epids=epids)
    sl = slice(0, npv)
    return om.evaluate(wl.k[sl], wl.ldc[sl], wl.t0[sl], wl.p[sl], wl.a[sl], wl.i[sl], wl.e[sl], wl.w[sl])


def gpu_flux(wl, precision, npv=None, lcids=None, pbids=None, epids=None, **kw):
    tm = RoadRunnerModel(wl.law, precision=precision, **kw)
    tm.set_data(wl.times, lcids=lcids, pbids=pbids, nsamples=wl.nsamples, exptimes=wl.exptimes, epids=epids)
    sl = slice(0, npv)
    f = tm.evaluate(wl.k[sl], wl.ldc[sl], wl.t0[sl], wl.p[sl], wl.a[sl], wl.i[sl], wl.e[sl], wl.w[sl])
    return np.asarray(f, np.float64).reshape(-1, len(wl.times))


def test_weight_table(gpu, oracle):
    say('case: weight table 256x100x40 fp64')
    tm = RoadRunnerModel('quadratic')
    W = tm.weights
    dk, dg, Wo = oracle.weights_3d(256, 0.005, 0.5, tm.ze, 100)
    assert tm.dk == dk and tm.dg == dg
    # Table entries are differences of lens areas, so the 1-ulp differences of the
    # GPU and host acos are amplified by the cancellation; 1e-11 is far inside what
    # the 1e-9 flux bar needs.
    err = float(np.abs(W - Wo).max())
    assert err < 1e-11, err


@pytest.mark.parametrize('law,coef', [
    ('uniform', [0.0, 0.0]), ('linear', [0.3]), ('quadratic', [0.3, 0.2]), ('quadratic-tri', [0.5, 0.3]),
    ('nonlinear', [0.1, 0.2, 0.1, 0.05]), ('general', [0.5, 0.3, 0.1]), ('square_root', [0.3, 0.2]),
    ('power-2', [0.6, 0.7]), ('power-2-pm', [0.6, 0.5]), ('logarithmic', [0.3, 0.2]), ('exponential', [0.3, 0.01])])
def test_ld_profiles(gpu, oracle, law, coef):
    say('case: ld profiles', law)
    lib = _lib.load()
    ze, zm = oracle.create_z_grid(0.7, 20, 20)
    mu = oracle.mu_nodes(zm)
    rng = np.random.default_rng(5)
    ldc = np.asarray(coef)[None, None, :] * (1 + 0.1 * rng.random((3, 2, len(coef))))
    ldp_o, ist_o = oracle.evaluate_ld(law, mu, ldc)
    c = torch.as_tensor(ldc, device=gpu)
    m = torch.as_tensor(mu, device=gpu)
    ldp = torch.empty((3, 2, mu.size), dtype=torch.float64, device=gpu)
    ist = torch.empty((3, 2), dtype=torch.float64, device=gpu)
    _lib.check(lib.rr_ld_profiles(_lib.LD_LAW_IDS[law][0], c.data_ptr(), 3, 2, len(coef), m.data_ptr(), mu.size,
                                  ldp.data_ptr(), ist.data_ptr(), None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(ldp.cpu().numpy(), ldp_o, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(ist.cpu().numpy(), ist_o, rtol=1e-12)


def test_orbit_coefficients_and_window(gpu, oracle):
    say('case: orbit coefficients, 64 pv incl. eccentric')
    rng = np.random.default_rng(11)
    n = 64
    p = rng.uniform(1, 13, n)
    a = workloads.as_from_rhop(rng.uniform(0.5, 3, n), p)
    inc = workloads.i_from_ba(rng.uniform(0, 0.9, n), a)
    e = np.where(np.arange(n) % 2 == 0, 0.0, rng.uniform(0, 0.5, n))
    w = rng.uniform(0, 2 * np.pi, n)
    k0 = rng.uniform(0.05, 0.15, n)
    lib = _lib.load()
    dev = [torch.as_tensor(x, device=gpu) for x in (p, a, inc, e, w, k0)]
    co = torch.empty((n, 2, 5), dtype=torch.float64, device=gpu)
    win = torch.empty((n, 2), dtype=torch.float64, device=gpu)
    _lib.check(lib.rr_orbit(n, *[x.data_ptr() for x in dev], co.data_ptr(), win.data_ptr(), None))
    torch.cuda.synchronize()
    co = co.cpu().numpy()
    win = win.cpu().numpy()
    tt = np.linspace(-0.3, 0.3, 601)
    for j in range(n):
        c = oracle.taylor_coeffs(p[j], a[j], inc[j], e[j], w[j])
        # The n-th stencil divides positions of size ~a by dt^n (dt = 0.02 d), so
        # 1-ulp differences of sin/cos grow to ~40 a eps / dt^n in coefficient n.
        bound = 40 * a[j] * np.finfo(float).eps / 0.02 ** np.arange(5) + 1e-15
        assert (np.abs(co[j] - c) <= bound[None, :]).all(), (j, co[j] - c)
        zbound = 2 * (bound[None, :] * np.abs(tt)[:, None] ** np.arange(5)).sum(1)
        assert (np.abs(oracle.sep(tt, co[j]) - oracle.sep(tt, c)) <= zbound).all()
        bt = oracle.bounding_box(k0[j], c)
        np.testing.assert_allclose(win[j], bt, atol=1e-9)


def test_rr_small_fp64(gpu, oracle):
    say('case: rr_small 64x2000 quadratic fp64 vs oracle')
    wl = workloads.rr_small()
    fo = oracle_flux(oracle, wl, 'quadratic')
    fg = gpu_flux(wl, 'fp64')
    assert np.array_equal(np.isnan(fo), np.isnan(fg))
    assert np.nanmax(np.abs(fg - fo)) <= FP64_TOL
    assert (fo < 1).sum() > 1000  # the case has transits


def test_rr_small_mixed(gpu, oracle):
    say('case: rr_small 64x2000 quadratic mixed vs oracle')
    wl = workloads.rr_small()
    fo = oracle_flux(oracle, wl, 'quadratic')
    fg = gpu_flux(wl, 'mixed')
    err = np.nanmax(np.abs(fg - fo))
    say(f'  max |dF| = {err * 1e6:.4f} ppm')
    assert err <= MIXED_TOL


@pytest.mark.parametrize('npv', [300, 257, 1])
def test_tile_groups_with_a_partial_last_group(gpu, oracle, npv):
    # The flux kernel walks the pvs in groups of 256 (within a group the same time
    # tile of every pv first), dividing a block index by the group's size with a
    # host-computed multiplier: 300 pv x 3000 t (3 tiles) is one full group and one
    # of 44; 257 pv a full group and one of a single pv (divisor 1); 1 pv one group
    # of one.
    say(f'case: {npv} pv x 3000 t, nss=3, quadratic, fp64 and mixed (tile groups of 256 and {npv % 256})')
    wl = workloads.rr_small(npv=npv, npt=3000, seed=7)
    wl.nsamples, wl.exptimes = 3, 0.02
    fo = oracle_flux(oracle, wl, 'quadratic')
    assert (fo < 1).sum() > 30 * npv
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


@pytest.mark.parametrize('nss', [25, 64])
def test_many_samples_small_chunks(gpu, oracle, nss):
    # 1024 (point, sample) items per phase-B chunk: nss = 25 gives chunks of 40
    # points, nss = 64 of 16 (the FP32 small-planet loop's limit), so a window's
    # points span several chunks of fewer than 64
    say(f'case: 40 pv x 3000 t, nss={nss} over 30 min, quadratic, fp64 and mixed')
    wl = workloads.rr_small(npv=40, npt=3000, seed=9)
    wl.nsamples, wl.exptimes = nss, 30.0 / 1440.0
    fo = oracle_flux(oracle, wl, 'quadratic')
    assert (fo < 1).sum() > 40 * 30
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


@pytest.mark.parametrize('stress', [False, True])
def test_rr_bjd_slice(gpu, oracle, stress):
    say(f'case: rr_bjd slice 24x20000 power-2 nss=10 stress={stress}')
    wl = workloads.rr_bjd(stress=stress)
    npv = 24
    fo = oracle_flux(oracle, wl, 'power-2', npv=npv)
    f64 = gpu_flux(wl, 'fp64', npv=npv)
    fmx = gpu_flux(wl, 'mixed', npv=npv)
    e64 = np.nanmax(np.abs(f64 - fo))
    emx = np.nanmax(np.abs(fmx - fo))
    say(f'  fp64 max |dF| = {e64:.3e}, mixed max |dF| = {emx * 1e6:.4f} ppm, in-transit pts {(fo < 1).sum()}')
    assert (fo < 1).sum() > 0
    assert e64 <= FP64_TOL
    assert emx <= MIXED_TOL


def test_heterogeneous_light_curves(gpu, oracle):
    say('case: 3 light curves, 2 passbands, 2 epochs, per-lc supersampling')
    rng = np.random.default_rng(21)
    npt, npv = 3000, 16
    times = np.sort(rng.uniform(-0.3, 0.3, npt))
    lcids = rng.integers(0, 3, npt)
    pbids = np.array([0, 1, 1])
    epids = np.array([0, 1, 0])
    nsamples = np.array([1, 3, 5])
    exptimes = np.array([0.0, 0.01, 0.02])
    k = rng.uniform(0.08, 0.12, (npv, 2))
    t0 = rng.normal(0, 0.01, (npv, 2))
    p = rng.uniform(1, 5, npv)
    a = workloads.as_from_rhop(rng.uniform(1, 3, npv), p)
    inc = workloads.i_from_ba(rng.uniform(0, 0.8, npv), a)
    e = rng.uniform(0, 0.3, npv)
    w = rng.uniform(0, 2 * np.pi, npv)
    ldc = rng.uniform(0.1, 0.4, (npv, 2, 2))
    a[3] = 0.9          # invalid: a <= 1
    e[5] = -0.1         # invalid: e < 0
    ldc[7, 0, 0] = np.nan  # invalid: NaN LD
    om = oracle.OracleRR('quadratic')
    om.set_data(times, lcids=lcids, pbids=pbids, nsamples=nsamples, exptimes=exptimes, epids=epids)
    fo = om.evaluate(k, ldc, t0, p, a, inc, e, w)
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        tm = RoadRunnerModel('quadratic', precision=prec)
        tm.set_data(times, lcids=lcids, pbids=pbids, nsamples=nsamples, exptimes=exptimes, epids=epids)
        fg = np.asarray(tm.evaluate(k, ldc, t0, p, a, inc, e, w), float)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        assert np.isnan(fg[[3, 5, 7]]).all()
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


def test_contiguous_light_curves_run_classification(gpu, oracle):
    say('case: 3 contiguous light curves (same times per band), per-lc supersampling, ragged lengths')
    rng = np.random.default_rng(22)
    t_lc = [np.linspace(-0.3, 0.3, n) for n in (997, 1003, 1500)]
    times = np.concatenate(t_lc)
    lcids = np.concatenate([np.full(t.size, j) for j, t in enumerate(t_lc)])
    pbids, epids = np.array([0, 1, 2]), np.array([0, 1, 0])
    nsamples, exptimes = np.array([1, 4, 7]), np.array([0.0, 0.01, 0.02])
    npv = 24
    k = rng.uniform(0.08, 0.12, (npv, 3))
    t0 = rng.normal(0, 0.01, (npv, 2))
    p = rng.uniform(1, 5, npv)
    a = workloads.as_from_rhop(rng.uniform(1, 3, npv), p)
    inc = workloads.i_from_ba(rng.uniform(0, 0.8, npv), a)
    e, w = np.zeros(npv), np.zeros(npv)
    ldc = rng.uniform(0.1, 0.4, (npv, 3, 2))
    a[4] = 0.9
    om = oracle.OracleRR('quadratic')
    om.set_data(times, lcids=lcids, pbids=pbids, nsamples=nsamples, exptimes=exptimes, epids=epids)
    fo = om.evaluate(k, ldc, t0, p, a, inc, e, w)
    assert (fo < 1).sum() > 5000
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        tm = RoadRunnerModel('quadratic', precision=prec)
        tm.set_data(times, lcids=lcids, pbids=pbids, nsamples=nsamples, exptimes=exptimes, epids=epids)
        assert tm._device_data(tm.device)['sorted']
        fg = np.asarray(tm.evaluate(k, ldc, t0, p, a, inc, e, w), float)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


def test_bjd_segments_many_passbands(gpu, oracle):
    say('case: BJD 2-min cadence split into 11 contiguous passband light curves, nss=5, power-2, k0 off-table')
    wl = workloads.rr_bjd(npv=32, npt=6000, seed=9)
    wl.nsamples = 5
    npb = 11  # two limb-darkening groups in the setup kernel (8 + 3)
    lcids = (np.arange(wl.times.size) * npb) // wl.times.size
    rng = np.random.default_rng(9)
    wl.k = wl.k + 0.002 * np.sin(np.arange(npb) / 3.0)[None, :]
    wl.k[0, 0] = 0.52  # outside klims: exact 2-D weights for every passband of pv 0
    wl.ldc = np.concatenate([wl.ldc] * npb, 1) + rng.uniform(-0.05, 0.05, (32, npb, 2))
    pbids = np.arange(npb)
    fo = oracle_flux(oracle, wl, 'power-2', lcids=lcids, pbids=pbids)
    assert (fo < 1).sum() > 1000
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec, lcids=lcids, pbids=pbids)
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


@pytest.mark.parametrize('law,npb,nldc', [('quadratic', 9, 2), ('nonlinear', 9, 4), ('square_root', 7, 2),
                                          ('general', 6, 3), ('linear', 5, 1), ('quadratic-tri', 6, 2)])
def test_linear_laws_many_passbands_basis(gpu, oracle, law, npb, nldc):
    # more passbands than basis functions: the setup evaluates the limb-darkening
    # means and disk integrals through the law's basis (one table pass per pv);
    # pv 0 has k off the table (exact 2-D weights), pv 1 a NaN coefficient in
    # passband 0 (a NaN row), pv 2 a NaN coefficient in passband 3
    say(f'case: {law}, {npb} contiguous passband light curves, nss=3, basis limb darkening')
    wl = workloads.rr_bjd(npv=20, npt=3000, seed=11)
    wl.law = law
    wl.nsamples = 3
    lcids = (np.arange(wl.times.size) * npb) // wl.times.size
    rng = np.random.default_rng(11)
    wl.k = wl.k + 0.002 * np.sin(np.arange(npb) / 3.0)[None, :]
    wl.k[0, 0] = 0.52
    base = {1: [0.5], 2: [0.4, 0.2], 3: [0.5, -0.1, 0.05], 4: [0.56, -0.08, 0.67, -0.33]}[nldc]
    wl.ldc = np.asarray(base)[None, None, :] + rng.uniform(-0.03, 0.03, (20, npb, nldc))
    wl.ldc[1, 0, 0] = np.nan
    wl.ldc[2, 3, -1] = np.nan
    pbids = np.arange(npb)
    fo = oracle_flux(oracle, wl, law, lcids=lcids, pbids=pbids)
    assert (fo < 1).sum() > 500
    assert np.isnan(fo[1]).all()
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec, lcids=lcids, pbids=pbids)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


@pytest.mark.parametrize('nzin,nzlimb,npb', [(64, 64, 9), (400, 400, 1), (400, 400, 3)])
def test_setup_lds_layout_large_z_grids(gpu, oracle, nzin, nzlimb, npb):
    # The setup block's LDS is sized per launch.  nz = 128 with 9 passbands fills
    # the profile area (8 x 128 doubles) and leaves a last group of one passband
    # on the 16-byte path; nz = 800 leaves room for one profile plus that path's
    # two row sums (nz + 2 ng = 1000 of 1024 doubles), with the disk integrals
    # and the 200 trapezoid nodes after them.
    say(f'case: power-2, nz = {nzin + nzlimb}, {npb} passband(s), nss=3, setup LDS layout')
    npv = 12
    wl = workloads.rr_bjd(npv=npv, npt=2000, seed=13)
    wl.nsamples = 3
    lcids = pbids = None
    if npb > 1:
        rng = np.random.default_rng(13)
        lcids = (np.arange(wl.times.size) * npb) // wl.times.size
        pbids = np.arange(npb)
        wl.k = wl.k + 0.002 * np.sin(np.arange(npb) / 3.0)[None, :]
        wl.ldc = wl.ldc[:, :1, :] + rng.uniform(-0.03, 0.03, (npv, npb, 2))
    fo = oracle_flux(oracle, wl, 'power-2', nzin=nzin, nzlimb=nzlimb, lcids=lcids, pbids=pbids)
    assert (fo < 1).sum() > 300
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec, lcids=lcids, pbids=pbids, nzin=nzin, nzlimb=nzlimb)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


def test_short_periods_window_search(gpu, oracle):
    # periods of 0.45-1.6 d against 1.42-d tiles (1024 times at 2-min cadence):
    # tiles spanning one, two (window ranges found by search) and three or more
    # epochs (run classification), windows across tile edges, a 30-min exposure
    say('case: 24 pv, periods 0.45-1.6 d, 6000 times at 2-min cadence, nss=7')
    rng = np.random.default_rng(31)
    npv = 24
    wl = workloads.rr_bjd(npv=npv, npt=6000, seed=31)
    wl.nsamples = 7
    wl.p = rng.uniform(0.45, 1.6, npv)
    wl.a = workloads.as_from_rhop(rng.uniform(1.0, 3.0, npv), wl.p)
    wl.i = workloads.i_from_ba(rng.uniform(0.0, 0.9, npv), wl.a)
    wl.t0 = (wl.times[0] + rng.uniform(0.0, 1.0, npv))[:, None]
    fo = oracle_flux(oracle, wl, 'power-2')
    assert (fo < 1).sum() > 5000
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


@pytest.mark.parametrize('case', ['bjd', 'fast'])
def test_samples_beyond_contacts(gpu, oracle, case):
    # One light curve with one nsamples.  A point in the window whose samples all
    # lie outside the disk, like a point outside the window, must come out exactly
    # 1 in both paths wherever the oracle gives exactly 1 (the FP32 depth sum is 0,
    # the FP64 sum is ns ones).  'fast': periods of 0.3-0.6 d with a 1-h exposure,
    # where a sample reaches 0.09 d past a contact.
    rng = np.random.default_rng(5)
    npv = 32
    wl = workloads.rr_bjd(npv=npv, npt=6000, seed=5)
    if case == 'fast':
        wl.nsamples, wl.exptimes = 4, 1.0 / 24
        wl.p = rng.uniform(0.3, 0.6, npv)
        wl.a = workloads.as_from_rhop(rng.uniform(1.0, 3.0, npv), wl.p)
        wl.i = workloads.i_from_ba(rng.uniform(0.0, 0.95, npv), wl.a)
        wl.t0 = (wl.times[0] + rng.uniform(0.0, 1.0, npv))[:, None]
    say(f'case: samples beyond the contacts, {case}: {npv} pv x 6000 t, nss={wl.nsamples}')
    fo = oracle_flux(oracle, wl, 'power-2')
    ones = fo == 1.0
    assert ones.sum() > 1000 and (fo < 1).sum() > 1000
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}, non-ones where the oracle has ones: {(fg[ones] != 1.0).sum()}')
        assert (fg[ones] == 1.0).all()
        assert err <= tol


def test_k_outside_table_uses_exact_weights(gpu, oracle):
    say('case: k outside klims -> exact 2-D weights')
    wl = workloads.rr_small(npv=8)
    wl.k = wl.k.copy()
    wl.k[:4, 0] = [0.003, 0.0045, 0.55, 0.6]
    fo = oracle_flux(oracle, wl, 'quadratic')
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


def test_tabulated_profiles_reproduce_the_law(gpu, oracle):
    # Profiles given as a table at the model's mu nodes (an LDModel), here the
    # quadratic law's own values with its analytic disk integral, give the
    # quadratic model: the tabulated-profile branch of the setup against the oracle.
    say('case: tabulated LD profiles (quadratic values at the mu nodes), one passband')
    wl = workloads.rr_small(npv=8)
    wl.k = wl.k.copy()
    wl.k[0, 0] = 0.55  # off the table: exact 2-D weights on tabulated profiles
    fo = oracle_flux(oracle, wl, 'quadratic')
    c = np.asarray(wl.ldc, np.float64).reshape(wl.k.shape[0], 1, 2)
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        probe = RoadRunnerModel('quadratic', precision=prec)
        mu = np.asarray(probe.mu, np.float64)
        ldp = np.stack([oracle.ld_quadratic(mu, c[j, 0]) for j in range(c.shape[0])])[:, None, :]
        istar = np.array([[oracle.ldi_quadratic(c[j, 0])] for j in range(c.shape[0])])
        tm = RoadRunnerModel(TabulatedLDModel(ldp, istar), precision=prec)
        tm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
        fg = np.asarray(tm.evaluate(wl.k, wl.ldc, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w), np.float64)
        fg = fg.reshape(-1, len(wl.times))
        assert np.array_equal(np.isnan(fo), np.isnan(fg)), prec
        err = np.nanmax(np.abs(fg - fo))
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


def test_single_pv_scalars(gpu, oracle):
    say('case: scalar parameters (rr_simple path)')
    times = np.linspace(-0.2, 0.2, 500)
    tm = RoadRunnerModel('quadratic')
    tm.set_data(times)
    f = tm.evaluate(0.1, [0.3, 0.1], 0.0, 3.0, 8.0, 0.5 * math.pi)
    om = oracle.OracleRR('quadratic')
    om.set_data(times)
    fo = om.evaluate(np.array([[0.1]]), np.array([0.3, 0.1]), np.array([[0.0]]), 3.0, 8.0, 0.5 * math.pi)[0]
    assert f.shape == (500,)
    assert np.abs(f - fo).max() <= FP64_TOL


def test_empty_and_device_output(gpu):
    say('case: zero times; copy=False returns a device tensor')
    tm = RoadRunnerModel('quadratic')
    tm.set_data(np.linspace(-0.1, 0.1, 64))
    out = tm.evaluate(np.full(4, 0.1), np.tile([[[0.3, 0.1]]], (4, 1, 1)), np.zeros(4), np.full(4, 3.0),
                      np.full(4, 8.0), np.full(4, 1.5), copy=False)
    assert isinstance(out, torch.Tensor) and out.is_cuda and out.shape == (4, 64)


def test_unsorted_times(gpu, oracle):
    say('case: rr_small with shuffled times (general flux path)')
    wl = workloads.rr_small(npv=16)
    perm = np.random.default_rng(3).permutation(wl.times.size)
    wl.times = wl.times[perm]
    fo = oracle_flux(oracle, wl, 'quadratic')
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        assert np.nanmax(np.abs(fg - fo)) <= tol, prec


def test_npt_not_multiple_of_run(gpu, oracle):
    say('case: npt = 1001 (ragged tail, unaligned rows)')
    wl = workloads.rr_small(npv=7, npt=1001)
    fo = oracle_flux(oracle, wl, 'quadratic')
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        assert np.nanmax(np.abs(fg - fo)) <= tol, prec


def test_all_points_in_transit_overflow_path(gpu, oracle):
    say('case: 64 pv x 2000 times all in transit (queue overflow computed in place)')
    wl = workloads.rr_small(npv=64, npt=2000)
    wl.times = np.linspace(-0.02, 0.02, 2000)
    wl.nsamples, wl.exptimes = 3, 0.005
    fo = oracle_flux(oracle, wl, 'quadratic')
    assert (fo < 1).mean() > 0.9
    for prec, tol in (('fp64', FP64_TOL), ('mixed', MIXED_TOL)):
        fg = gpu_flux(wl, prec)
        assert np.nanmax(np.abs(fg - fo)) <= tol, prec
