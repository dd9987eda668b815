"""Seeded synthetic inputs for the benchmark and parity workloads.

Each generator draws, with ``numpy.random.default_rng(seed)``, parameter
populations in the layouts ``evaluate`` takes.  ``a`` comes from the stellar
density through the scaled semi-major axis relation (pytransit/orbits/orbits_py.py:
603-618) and ``i`` from the impact parameter (orbits_py.py:673-688).

* ``rr_small``      -- 64 pv x 2,000 times over [-0.5, 0.5] d, quadratic LD, no
                       supersampling (the CPU-runnable case).
* ``rr_bjd``        -- 4,096 pv x 20,000 times at 2-min cadence from BJD 2459000.0,
                       power-2 LD, 10 supersamples over 30 min.
* ``ts_many_pb``    -- 1,024 pv x 5,000 times x 512 passbands, tabulated 128-node
                       profiles (TSModel).
* ``rr_gradient``   -- 2,048 pv x 10,000 times, quadratic LD, 5 supersamples, with a
                       synthetic observed light curve.
* ``serving_stream`` -- Poisson arrivals of fit requests: light-curve sets of
                       log-uniform 1k-200k points in 1-64 passbands, populations of
                       log-uniform 32-8,192 pv, quadratic / power-2 / nonlinear-4
                       LD, nss in {1, 5, 10}.
"""
import math
from types import SimpleNamespace

import numpy as np

G_NEWTON = 6.6743e-11  # m^3 kg^-1 s^-2 (CODATA 2018)
SECONDS_PER_DAY = 86400.0


def as_from_rhop(rho, period):
    """Scaled semi-major axis a/R* from stellar density [g/cm^3] and period [d]."""
    rho = np.asarray(rho, float)
    period = np.asarray(period, float)
    return (G_NEWTON / (3 * math.pi)) ** (1 / 3) * ((period * SECONDS_PER_DAY) ** 2 * 1e3 * rho) ** (1 / 3)


def i_from_ba(b, a):
    """Inclination [rad] from impact parameter and scaled semi-major axis."""
    return np.arccos(np.asarray(b, float) / np.asarray(a, float))


def _orbit_draws(rng, npv, pmax):
    k = rng.uniform(0.05, 0.15, npv)
    p = rng.uniform(1.0, pmax, npv)
    rho = rng.uniform(0.5, 3.0, npv)
    b = rng.uniform(0.0, 0.9, npv)
    t0 = rng.normal(0.0, 0.005, npv)
    a = as_from_rhop(rho, p)
    inc = i_from_ba(b, a)
    return k, p, a, inc, t0


def rr_small(npv=64, npt=2000, seed=0):
    """64 x 2,000 points, quadratic LD, FP64 reference arithmetic."""
    rng = np.random.default_rng(seed)
    k, p, a, inc, t0 = _orbit_draws(rng, npv, 10.0)
    u1 = rng.uniform(0.1, 0.6, npv)
    u2 = rng.uniform(0.0, 0.3, npv)
    return SimpleNamespace(
        name='rr_small', law='quadratic', times=np.linspace(-0.5, 0.5, npt), nsamples=1, exptimes=0.0,
        k=k[:, None], t0=t0[:, None], p=p, a=a, i=inc, e=np.zeros(npv), w=np.zeros(npv),
        ldc=np.stack([u1, u2], 1)[:, None, :])


def rr_bjd(npv=4096, npt=20000, seed=1, stress=False):
    """4,096 x 20,000 points at BJD scale, power-2 LD, nss = 10 over 30 min.

    ``stress=True`` draws t0 around 0 instead of around the first time, so the
    epoch numbers reach ~2.5e6 and the FP64 epoch reduction is exercised.
    """
    rng = np.random.default_rng(seed)
    k, p, a, inc, t0 = _orbit_draws(rng, npv, 13.0)
    c = rng.uniform(0.3, 0.9, npv)
    alpha = rng.uniform(0.3, 1.0, npv)
    t_start = 2459000.0
    if not stress:
        t0 = t_start + t0
    return SimpleNamespace(
        name='rr_bjd' + ('_stress' if stress else ''), law='power-2',
        times=t_start + np.arange(npt) * (2.0 / 1440.0), nsamples=10, exptimes=30.0 / 1440.0,
        k=k[:, None], t0=t0[:, None], p=p, a=a, i=inc, e=np.zeros(npv), w=np.zeros(npv),
        ldc=np.stack([c, alpha], 1)[:, None, :])


def ts_many_pb(npv=1024, npt=5000, npb=512, nmu_in=64, nmu_limb=64, seed=2, mu=None):
    """TSModel case: per-passband k and tabulated quadratic-shaped profiles.

    ``mu`` are the model's 128 mu nodes (zcut 0.7, 64 + 64 annuli); the profile is
    ``1 - u1(1-mu) - u2(1-mu)^2`` with u1, u2 varying smoothly with passband plus
    N(0, 0.01) noise; istar is the analytic quadratic integral.
    """
    rng = np.random.default_rng(seed)
    k0 = rng.uniform(0.08, 0.12, npv)
    p = rng.uniform(1.0, 10.0, npv)
    rho = rng.uniform(0.5, 3.0, npv)
    b = rng.uniform(0.0, 0.9, npv)
    t0 = rng.normal(0.0, 0.005, npv)
    a = as_from_rhop(rho, p)
    inc = i_from_ba(b, a)
    j = np.arange(npb)
    k = k0[:, None] + 0.002 * np.sin(j / 40.0)[None, :]
    u1 = (0.55 - 0.25 * j / 511.0)[None, :] + rng.normal(0.0, 0.01, (npv, npb))
    u2 = (0.25 - 0.10 * j / 511.0)[None, :] + rng.normal(0.0, 0.01, (npv, npb))
    ns = SimpleNamespace(name='ts_many_pb', times=np.linspace(-0.2, 0.2, npt), nsamples=1, exptimes=0.0,
                         k=k, t0=t0, p=p, a=a, i=inc, e=np.zeros(npv), w=np.zeros(npv), u1=u1, u2=u2,
                         nzin=nmu_in, nzlimb=nmu_limb)
    if mu is not None:
        x = 1.0 - np.asarray(mu)[None, None, :]
        ns.ldp = 1.0 - u1[..., None] * x - u2[..., None] * x ** 2
        ns.istar = 2 * math.pi / 12 * (-2 * u1 - u2 + 6)
    return ns


def rr_gradient(npv=2048, npt=10000, seed=3, exptime=0.01, noise=300e-6):
    """Likelihood/gradient case: a true parameter set, a noisy synthetic light curve
    (added by the caller from a model evaluation) and a population around the truth."""
    rng = np.random.default_rng(seed)
    k, p, a, inc, t0 = _orbit_draws(rng, 1, 10.0)
    u1 = rng.uniform(0.1, 0.6, 1)
    u2 = rng.uniform(0.0, 0.3, 1)
    truth = dict(k=k[0], t0=t0[0], p=p[0], a=a[0], i=inc[0], u1=u1[0], u2=u2[0])
    sig = dict(k=2e-3, t0=1e-3, p=1e-4, a=0.05, i=2e-3, u1=0.02, u2=0.02)
    pop = {n: truth[n] + rng.normal(0.0, sig[n], npv) for n in truth}
    pop['i'] = np.minimum(pop['i'], 0.5 * math.pi)
    return SimpleNamespace(name='rr_gradient', law='quadratic', times=np.linspace(-0.35, 0.35, npt), nsamples=5,
                           exptimes=exptime, truth=truth, pop=pop, noise=noise,
                           noise_draw=rng.normal(0.0, noise, npt))


LAW_DRAWS = ('quadratic', 'power-2', 'nonlinear')


def _ld_draw(rng, law, npv, npb):
    """Limb-darkening coefficients [npv, npb, nldc] for one request."""
    if law == 'quadratic':
        return np.stack([rng.uniform(0.1, 0.6, (npv, npb)), rng.uniform(0.0, 0.3, (npv, npb))], -1)
    if law == 'power-2':
        return np.stack([rng.uniform(0.3, 0.9, (npv, npb)), rng.uniform(0.3, 1.0, (npv, npb))], -1)
    # four-parameter nonlinear law near solar-type values
    base = np.array([0.56, -0.08, 0.67, -0.33])
    return base + rng.normal(0.0, 0.02, (npv, npb, 4))


def serving_dataset(rng, npt_range=(1000, 200000), npb_max=64):
    """One light-curve set: npt ~ log-uniform, 2-min cadence from BJD 2459000.0,
    split into npb contiguous light curves, one per passband; nss in {1, 5, 10}
    with a 30-min exposure when nss > 1."""
    npt = int(round(math.exp(rng.uniform(math.log(npt_range[0]), math.log(npt_range[1])))))
    npb = int(rng.integers(1, npb_max + 1))
    npb = min(npb, npt)
    nss = int(rng.choice([1, 5, 10]))
    times = 2459000.0 + np.arange(npt) * (2.0 / 1440.0)
    lcids = (np.arange(npt) * npb) // npt
    return dict(time=times, lcids=lcids, pbids=np.arange(npb), nsamples=nss,
                exptimes=(30.0 / 1440.0) if nss > 1 else 0.0)


def serving_request(rng, npb, npv_range=(32, 8192)):
    """One fit request's population on a dataset with npb passbands."""
    npv = int(round(math.exp(rng.uniform(math.log(npv_range[0]), math.log(npv_range[1])))))
    law = str(rng.choice(LAW_DRAWS))
    k, p, a, inc, t0 = _orbit_draws(rng, npv, 13.0)
    kpb = k[:, None] + 0.002 * np.sin(np.arange(npb) / 8.0)[None, :]
    return dict(k=kpb, ldc=_ld_draw(rng, law, npv, npb), t0=2459000.0 + t0, p=p, a=a, i=inc, ld=law)


def serving_stream(n_requests=300, rate=2000.0, n_datasets=24, seed=4, npt_range=(1000, 200000),
                   npv_range=(32, 8192), npb_max=64):
    """A seeded request stream: ``datasets`` (set_data keyword dicts) and
    ``requests`` (submit keyword dicts with ``dataset`` index and ``arrival`` in s
    from the start, Poisson at ``rate`` per second).  Each request draws its
    light-curve set uniformly from the pool, so sets are reused and stay resident."""
    rng = np.random.default_rng(seed)
    datasets = [serving_dataset(rng, npt_range, npb_max) for _ in range(n_datasets)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    requests = []
    for j in range(n_requests):
        d = int(rng.integers(n_datasets))
        r = serving_request(rng, datasets[d]['pbids'].size, npv_range)
        r.update(dataset=d, arrival=float(arrivals[j]))
        requests.append(r)
    return SimpleNamespace(name='serving_stream', datasets=datasets, requests=requests, rate=rate)
