"""GPU parity of the LDTk-style profile path: trilinear interpolation of a grid of
stellar intensity profiles over (teff, logg, metallicity) and its disk integrals
(numba/ldtkldm.py:5-91, LDTkLDModel.__call__ of models/ldtkldm.py:74-89) against
the oracle, and RoadRunnerModel and TSModel driven by it.  LDTk itself is not installed: the
grids here are synthetic, of the shape LDTk gives ([nteff, nlogg, nmetal, npb, nmu]
at the model's mu nodes).  Bars: profiles and integrals to FP64 rounding; fluxes
FP64 <= 1e-9, mixed <= 1 ppm."""
import numpy as np
import pytest
import torch

from pytransit_amd import LDTkLDModel, RoadRunnerModel, TabulatedLDModel, workloads

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_TOL = 1e-6

TEFFS = 4000.0 + 250.0 * np.arange(5)
LOGGS = 3.5 + 0.5 * np.arange(4)
METALS = -0.5 + 0.25 * np.arange(3)
AXES = (TEFFS[0], TEFFS[1] - TEFFS[0], LOGGS[0], LOGGS[1] - LOGGS[0], METALS[0], METALS[1] - METALS[0])


def say(*a):
    print(*a, flush=True)


def quadratic_grid(oracle, mu, npb):
    """Quadratic-law profiles whose coefficients vary smoothly over the grid."""
    g = np.empty((TEFFS.size, LOGGS.size, METALS.size, npb, mu.size))
    for i, t in enumerate(TEFFS):
        for j, lg in enumerate(LOGGS):
            for k, m in enumerate(METALS):
                for b in range(npb):
                    u1 = 0.6 - 0.05 * (t - 4000.0) / 250.0 + 0.02 * lg + 0.03 * m - 0.04 * b
                    u2 = 0.15 + 0.01 * (t - 4000.0) / 250.0 - 0.01 * lg + 0.02 * b
                    g[i, j, k, b] = oracle.ld_quadratic(mu, [u1, u2])
    return g


def stellar_params(rng, n):
    return np.stack([rng.uniform(TEFFS[0], TEFFS[-1], n), rng.uniform(LOGGS[0], LOGGS[-1], n),
                     rng.uniform(METALS[0], METALS[-1], n)], 1)


def test_ldtk_interpolation_and_integrals(gpu, oracle):
    say('case: rr_ldtk_interpolate, grid 5x4x3 x 2 passbands x 40 mu, 64 pv with clamped and NaN parameters')
    rng = np.random.default_rng(7)
    mu = np.asarray(RoadRunnerModel('quadratic').mu, np.float64)
    grid = rng.uniform(0.2, 1.0, (TEFFS.size, LOGGS.size, METALS.size, 2, mu.size))
    par = stellar_params(rng, 64)
    par[:4] = [[3000.0, 4.0, 0.0], [9000.0, 9.0, 1.0], [TEFFS[-1], LOGGS[-1], METALS[-1]], [4100.0, 3.0, -2.0]]
    par[5, 1] = np.nan
    m = LDTkLDModel(profiles=grid, teffs=TEFFS, loggs=LOGGS, metals=METALS)
    ldp, istar = m(mu, par)
    assert ldp.is_cuda and tuple(ldp.shape) == (64, 2, mu.size) and tuple(istar.shape) == (64, 2)
    ldp, istar = ldp.cpu().numpy(), istar.cpu().numpy()
    ldp_o = oracle.trilinear_interpolation_set(grid, par[:, 0], par[:, 1], par[:, 2], *AXES)
    ist_o = oracle.integrate_profiles(mu, ldp_o)
    assert np.array_equal(np.isnan(ldp), np.isnan(ldp_o)) and np.isnan(ldp[5]).all()
    e1 = np.nanmax(np.abs(ldp - ldp_o))
    e2 = np.nanmax(np.abs(istar - ist_o))
    say(f'  max |d ldp| = {e1:.2e}, max |d istar| = {e2:.2e}')
    assert e1 <= 1e-14 and e2 <= 1e-13
    # device parameters in, same result
    ldp2, _ = m(mu, torch.as_tensor(par, device=gpu))
    assert np.array_equal(ldp2.cpu().numpy(), ldp, equal_nan=True)


def test_roadrunner_with_ldtk_profiles(gpu, oracle):
    # RoadRunnerModel with the LDTk-style model gives the flux of the same profiles
    # handed over as a table (itself pinned against the oracle's quadratic model)
    say('case: RoadRunnerModel(LDTkLDModel) vs tabulated oracle-interpolated profiles, 8 pv x 2000 t')
    rng = np.random.default_rng(9)
    wl = workloads.rr_small(npv=8)
    mu = np.asarray(RoadRunnerModel('quadratic').mu, np.float64)
    grid = quadratic_grid(oracle, mu, 1)
    par = stellar_params(rng, 8)
    ldp_o = oracle.trilinear_interpolation_set(grid, par[:, 0], par[:, 1], par[:, 2], *AXES)
    ist_o = oracle.integrate_profiles(mu, ldp_o)
    for prec, tol in (('fp64', 1e-12), ('mixed', MIXED_TOL)):
        fs = []
        for ldm in (LDTkLDModel(profiles=grid, teffs=TEFFS, loggs=LOGGS, metals=METALS),
                    TabulatedLDModel(ldp_o, ist_o)):
            tm = RoadRunnerModel(ldm, precision=prec)
            tm.set_data(wl.times)
            f = tm.evaluate(wl.k, par, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w)
            fs.append(np.asarray(f, np.float64).reshape(-1, wl.times.size))
        assert (fs[1] < 1).sum() > 500
        err = np.abs(fs[0] - fs[1]).max()
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol


def test_tsmodel_with_ldtk_profiles(gpu, oracle):
    # the many-passband use LDTk grids are made for: TSModel with 16 passbands on
    # 128 mu nodes, against the same profiles handed over as a table
    from pytransit_amd import TSModel
    say('case: TSModel(LDTkLDModel), 4 pv x 16 pb x 1500 t, 128 mu nodes')
    rng = np.random.default_rng(11)
    npv, npb = 4, 16
    times = np.linspace(-0.3, 0.3, 1500)
    mu = np.asarray(TSModel('quadratic', nzin=64, nzlimb=64).mu, np.float64)
    grid = quadratic_grid(oracle, mu, npb)
    par = stellar_params(rng, npv)
    ldp_o = oracle.trilinear_interpolation_set(grid, par[:, 0], par[:, 1], par[:, 2], *AXES)
    ist_o = oracle.integrate_profiles(mu, ldp_o)
    k = 0.1 + 0.004 * np.sin(np.arange(npb))[None, :] + rng.uniform(-0.01, 0.01, (npv, 1))
    t0, p, a, inc = rng.uniform(-0.02, 0.02, npv), np.full(npv, 3.0), np.full(npv, 10.0), np.full(npv, 1.56)
    for prec, tol in (('fp64', 1e-12), ('mixed', MIXED_TOL)):
        fs = []
        for ldm in (LDTkLDModel(profiles=grid, teffs=TEFFS, loggs=LOGGS, metals=METALS),
                    TabulatedLDModel(ldp_o, ist_o)):
            tm = TSModel(ldm, nzin=64, nzlimb=64, precision=prec)
            tm.set_data(times)
            fs.append(np.asarray(tm.evaluate(k, par, t0, p, a, inc), np.float64))
        assert fs[0].shape == (npv, npb, times.size)
        assert (fs[1] < 1).sum() > 5000
        err = np.abs(fs[0] - fs[1]).max()
        say(f'  {prec}: max |dF| = {err:.3e}')
        assert err <= tol
