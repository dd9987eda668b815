"""FP64 numpy restatement of PyTransit's RoadRunner path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
it.  The product path (``pytransit_amd``) never calls it and fails loudly when
its HIP library is missing.

Every function cites the reference file:line it restates (paths relative to the
PyTransit 2.8.1 checkout).  All arithmetic is IEEE FP64, as in the reference.

Where the reference is undefined (out-of-bounds reads in numba with bounds
checking off) this restatement *defines* the behaviour:

* Q2  ``W[ik+1]`` past the k-table (``model_full.py:47-51``): ``ik`` is clamped
  to ``nk-2`` and ``ak`` is computed from the clamped ``ik`` (linear
  extrapolation from the last two rows).
* Q3  ``lda[ng]`` for g in (1-1e-7, 1] (``common.py:225-233``): the interval
  index is clamped to ``ng-2``.
* Q4/Q5 1-D ``k``/``t0`` of length npv are read as ``(npv, 1)``.

The orbit boundary (meepmeep ``solve2d``/``sep_c``/``bounding_box``) is not in
the reference checkout; it is restated from PyTransit's own in-repo Taylor-series
method (``orbits/taylor_z.py:23-102,219-226,298-328,391-394``).  The Taylor
coefficients are stored pre-divided by n! in a ``(2, 5)`` array (row 0: x,
row 1: y) and evaluated by Horner's rule; the x constant term is 0, as in
``taylor_z.py:224``.

Parity status of each part (how it is pinned) is listed in ``tests/test_oracle.py``.
"""
import math

import numpy as np

PI = math.pi
TWO_PI = 2.0 * math.pi
HALF_PI = 0.5 * math.pi

G_NEWTON = 6.6743e-11  # scipy.constants.G (CODATA 2018), as used by orbits_py.py:603-618
D_S = 24.0 * 60.0 * 60.0  # seconds per day, orbits_py.py:39-41


# --------------------------------------------------------------------------------------
# Radial grid and circle-circle geometry
# --------------------------------------------------------------------------------------

def create_z_grid(zcut: float, nin: int, nedge: int):
    """Annulus edges ``ze`` and means ``zm``; common.py:131-149.

    ``nin-1`` edges uniform in z up to (not including) zcut, then ``nedge+1`` edges
    uniform in mu from mu(zcut) down to 0 (z = 1).  ``zm[0]`` stays 0 (Q9).
    """
    n = nin + nedge
    mucut = math.sqrt(1.0 - zcut ** 2)
    dz = zcut / nin
    dmu = mucut / nedge
    ze = np.zeros(n)
    zm = np.zeros(n)
    for i in range(nin - 1):
        ze[i] = (i + 1) * dz
    for i in range(nedge + 1):
        ze[n - 1 - i] = math.sqrt(1.0 - (i * dmu) ** 2)
    for i in range(n - 1):
        zm[i + 1] = 0.5 * (ze[i] + ze[i + 1])
    return ze, zm


def mu_nodes(zm: np.ndarray) -> np.ndarray:
    """mu = sqrt(1 - zm^2); rrmodel.py:172."""
    return np.sqrt(1.0 - zm ** 2)


def cci_area(r1, r2, b):
    """Area of the intersection of two circles (radii r1, r2, separation b); common.py:36-49.

    Vectorised over broadcastable inputs; the branch order is the reference's.
    """
    r1, r2, b = np.broadcast_arrays(np.asarray(r1, float), np.asarray(r2, float), np.asarray(b, float))
    out = np.empty(r1.shape)
    c0 = r1 < b - r2
    c1 = ~c0 & (r1 >= b + r2)
    c2 = ~c0 & ~c1 & (b - r2 <= -r1)
    c3 = ~(c0 | c1 | c2)
    out[c0] = 0.0
    out[c1] = PI * r2[c1] ** 2
    out[c2] = PI * r1[c2] ** 2
    x1, x2, bb = r1[c3], r2[c3], b[c3]
    out[c3] = (x2 ** 2 * np.arccos((bb ** 2 + x2 ** 2 - x1 ** 2) / (2 * bb * x2))
               + x1 ** 2 * np.arccos((bb ** 2 + x1 ** 2 - x2 ** 2) / (2 * bb * x1))
               - 0.5 * np.sqrt((-bb + x2 + x1) * (bb + x2 - x1) * (bb - x2 + x1) * (bb + x2 + x1)))
    return out


def ccia_kite(r1, r2, b):
    """Lens area and kappa_0 by the kite formula of Agol et al. (2020); common.py:52-73 (+ tsort :5-33).

    Returns ``(area, kappa0)`` arrays.  The three lengths are sorted descending
    (``tsort``); sorting is order-independent for ties, so a generic sort is used.
    """
    r1, r2, b = np.broadcast_arrays(np.asarray(r1, float), np.asarray(r2, float), np.asarray(b, float))
    area = np.full(r1.shape, np.nan)
    kap = np.full(r1.shape, np.nan)
    none = r1 + r2 <= b
    lens = ~none & (np.abs(r1 - r2) < b) & (b <= r1 + r2)
    full2 = ~none & ~lens & (b <= r1 - r2)
    full1 = ~none & ~lens & ~full2 & (b <= r2 - r1)
    area[none] = 0.0
    kap[none] = 0.0
    area[full2] = PI * r2[full2] ** 2
    kap[full2] = PI
    area[full1] = PI * r1[full1] ** 2
    kap[full1] = 0.0
    if lens.any():
        s = np.sort(np.stack([r1[lens], r2[lens], b[lens]]), axis=0)
        z, y, x = s[0], s[1], s[2]
        akite = 0.5 * np.sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)))
        a1, a2, bb = r1[lens], r2[lens], b[lens]
        k0 = np.arctan2(2.0 * akite, (a2 - a1) * (a2 + a1) + bb * bb)
        k1 = np.arctan2(2.0 * akite, (a1 - a2) * (a1 + a2) + bb * bb)
        area[lens] = a1 * a1 * k1 + a2 * a2 * k0 - akite
        kap[lens] = k0
    return area, kap


def _seq_rowsum(w: np.ndarray) -> np.ndarray:
    """Left-to-right sum along the last axis (the reference's running ``s +=``)."""
    return np.cumsum(w, axis=-1)[..., -1]


def _linspace(a: float, b: float, n: int) -> np.ndarray:
    """numpy.linspace semantics (i*step + start, last element exactly b)."""
    return np.linspace(a, b, n)


def weights_2d(k: float, ze: np.ndarray, ng: int):
    """Normalised annulus-overlap weights W[ng, nz] for one radius ratio; common.py:152-185.

    Returns ``(gs, dg, W)``.
    """
    gs = _linspace(0.0, 1.0 - 1e-7, ng)
    b = gs * (1.0 + k)
    a = cci_area(ze[None, :], k, b[:, None])          # [ng, nz]
    w = np.empty_like(a)
    w[:, 0] = a[:, 0]
    w[:, 1:] = a[:, 1:] - a[:, :-1]
    w /= _seq_rowsum(w)[:, None]
    return gs, gs[1] - gs[0], w


def weights_3d(nk: int, kmin: float, kmax: float, ze: np.ndarray, ng: int):
    """3-D weight table W[nk, ng, nz] on ``linspace(kmin, kmax, nk)``; common.py:188-223.

    Returns ``(dk, dg, W)`` with **dk = (kmax-kmin)/nk** (Q1: the rows are built at a
    spacing of (kmax-kmin)/(nk-1) but indexed with this dk).
    """
    ks = _linspace(kmin, kmax, nk)
    gs = _linspace(0.0, 1.0 - 1e-7, ng)
    b = gs[None, :] * (1.0 + ks[:, None])              # [nk, ng]
    a = cci_area(ze[None, None, :], ks[:, None, None], b[:, :, None])
    w = np.empty_like(a)
    w[..., 0] = a[..., 0]
    w[..., 1:] = a[..., 1:] - a[..., :-1]
    w /= _seq_rowsum(w)[..., None]
    return (kmax - kmin) / nk, gs[1] - gs[0], w


def interp_ld_mean(g, dg: float, lda: np.ndarray):
    """Mean-limb-darkening interpolation; common.py:225-233 (Q3: index clamped to ng-2)."""
    g = np.asarray(g, float)
    ng = lda.shape[-1]
    out = np.empty(g.shape)
    neg = g < 0.0
    big = g > 1.0
    mid = ~neg & ~big
    gm = g[mid]
    i = np.floor(gm / dg).astype(np.int64)
    i = np.minimum(i, ng - 2)
    a = (gm - i * dg) / dg
    out[mid] = (1.0 - a) * lda[i] + a * lda[i + 1]
    out[neg] = np.nan
    out[big] = 0.0
    return out


# --------------------------------------------------------------------------------------
# Limb-darkening laws (numba/ldmodels.py:22-139) and batch evaluators (:142-175)
# --------------------------------------------------------------------------------------

def ld_uniform(mu, c):      # ldmodels.py:22-24
    return np.ones_like(mu)


def ld_linear(mu, c):       # ldmodels.py:32-34
    return 1.0 - c[0] * (1.0 - mu)


def ld_quadratic(mu, c):    # ldmodels.py:50-52
    return 1.0 - c[0] * (1.0 - mu) - c[1] * (1.0 - mu) ** 2


def ld_quadratic_tri(mu, c):  # ldmodels.py:74-78
    a, b = math.sqrt(c[0]), 2 * c[1]
    u, v = a * b, a * (1.0 - b)
    return 1.0 - u * (1.0 - mu) - v * (1.0 - mu) ** 2


def ld_nonlinear(mu, c):    # ldmodels.py:88-90
    return (1.0 - c[0] * (1.0 - np.sqrt(mu)) - c[1] * (1.0 - mu)
            - c[2] * (1.0 - np.power(mu, 1.5)) - c[3] * (1.0 - mu ** 2))


def ld_general(mu, c):      # ldmodels.py:93-98 (no leading 1: I = sum_i c_i (1 - mu^(i+1)))
    out = np.zeros_like(mu)
    for i in range(len(c)):
        out = out + c[i] * (1.0 - mu ** (i + 1))
    return out


def ld_square_root(mu, c):  # ldmodels.py:101-103
    return 1.0 - c[0] * (1.0 - mu) - c[1] * (1.0 - np.sqrt(mu))


def ld_logarithmic(mu, c):  # ldmodels.py:106-108 (IEEE semantics: mu=0 gives NaN)
    with np.errstate(divide='ignore', invalid='ignore'):
        return 1.0 - c[0] * (1.0 - mu) - c[1] * mu * np.log(mu)


def ld_exponential(mu, c):  # ldmodels.py:111-113 (IEEE semantics: mu=0 gives inf)
    with np.errstate(divide='ignore', invalid='ignore'):
        return 1.0 - c[0] * (1.0 - mu) - c[1] / (1.0 - np.exp(mu))


def ld_power_2(mu, c):      # ldmodels.py:116-118
    return 1.0 - c[0] * (1.0 - mu ** c[1])


def ld_power_2_pm(mu, c):   # ldmodels.py:135-139
    cc = 1 - c[0] + c[1]
    a = math.log2(cc / c[1])
    return 1.0 - cc * (1.0 - mu ** a)


def ldi_uniform(c):         # ldmodels.py:27-29
    return PI


def ldi_linear(c):          # ldmodels.py:37-39 (Q7: the reference's 2pi/6 (3 - 2u), replicated)
    return 2 * PI * 1 / 6 * (3 - 2 * c[0])


def ldi_quadratic(c):       # ldmodels.py:55-57
    return 2 * PI * 1 / 12 * (-2 * c[0] - c[1] + 6)


def ldi_quadratic_tri(c):   # ldmodels.py:81-85
    a, b = math.sqrt(c[0]), 2 * c[1]
    u, v = a * b, a * (1.0 - b)
    return 2 * PI * 1 / 12 * (-2 * u - v + 6)


def ldi_power_2(c):         # ldmodels.py:121-123
    return 2 * PI * (-c[0] * c[1] + c[1] + 2) / (2 * c[1] + 4)


# name -> (law, analytic integral or None); rrmodel.py:48-58
LD_LAWS = {
    'uniform': (ld_uniform, ldi_uniform),
    'linear': (ld_linear, ldi_linear),
    'quadratic': (ld_quadratic, ldi_quadratic),
    'quadratic-tri': (ld_quadratic_tri, ldi_quadratic_tri),
    'nonlinear': (ld_nonlinear, None),
    'general': (ld_general, None),
    'square_root': (ld_square_root, None),
    'logarithmic': (ld_logarithmic, None),
    'exponential': (ld_exponential, None),
    'power-2': (ld_power_2, ldi_power_2),
    'power-2-pm': (ld_power_2_pm, None),
}

_LDMU = np.linspace(1, 0, 200)           # rrmodel.py:151
_LDZ = np.sqrt(1 - _LDMU ** 2)           # rrmodel.py:152


def _ldc3(ldc) -> np.ndarray:
    """ldc reshape rules of evaluate_ld/evaluate_ldi; numba/ldmodels.py:144-150 (Q5)."""
    ldc = np.atleast_1d(np.asarray(ldc, float))
    if ldc.ndim == 1:
        return ldc.reshape((1, 1, -1))
    if ldc.ndim == 2:
        return ldc.reshape((1, ldc.shape[1], -1))
    return ldc


def _trapz(y, x) -> float:
    """scipy.integrate.trapezoid on 1-D data."""
    return float(np.sum((x[1:] - x[:-1]) * (y[1:] + y[:-1]) / 2.0))


def evaluate_ld(law: str, mu: np.ndarray, ldc):
    """ldp[npv, npb, nmu] and istar[npv, npb] for a named law.

    ldp: numba/ldmodels.py:142-157.  istar: the analytic integral
    (ldmodels.py:160-175) when the law has one, else the 200-node trapezoid
    2*pi*trapz(z*I, z) on mu = linspace(1, 0, 200) (rrmodel.py:220-227, Q8).
    """
    f, fi = LD_LAWS[law]
    c3 = _ldc3(ldc)
    npv, npb = c3.shape[:2]
    ldp = np.zeros((npv, npb, mu.size))
    istar = np.zeros((npv, npb))
    for ipv in range(npv):
        for ipb in range(npb):
            c = c3[ipv, ipb]
            ldp[ipv, ipb] = f(mu, c)
            if fi is not None:
                istar[ipv, ipb] = fi(c)
            else:
                istar[ipv, ipb] = 2 * PI * _trapz(_LDZ * f(_LDMU, c), _LDZ)
    return ldp, istar


def integrate_profiles(mu: np.ndarray, ldp: np.ndarray) -> np.ndarray:
    """Trapezoid-in-z disk integral of tabulated profiles; numba/ldtkldm.py:77-91."""
    z = np.sqrt(1.0 - mu ** 2)
    dz = z[1:] - z[:-1]
    s = dz * 0.5 * (z[1:] * ldp[..., 1:] + z[:-1] * ldp[..., :-1])
    return 2.0 * PI * np.cumsum(s, axis=-1)[..., -1]


def _grid_cell(x: float, x0: float, dx: float, n: int):
    """Cell index and fraction along one evenly spaced axis; numba/ldtkldm.py:40-51.
    The value is clamped to the axis and the index to n - 2, so the last node is
    reached with a fraction of 1.  (The reference clamps to x0 + n dx and then
    slices cell n - 1, which reads past its grid at and above the last node; inside
    the grid the two are the same.)"""
    x = min(max(x, x0), x0 + (n - 1) * dx)
    i = int(math.floor((x - x0) / dx))
    i = min(max(i, 0), n - 2)
    return i, (x - x0 - i * dx) / dx


def trilinear_interpolation_set(data, xs, ys, zs, x0, dx, y0, dy, z0, dz):
    """Profiles [npv, npb, nmu] interpolated trilinearly in data[nx, ny, nz, npb, nmu]
    at each (xs, ys, zs); numba/ldtkldm.py:5-57, with the corner weights and the
    sum in the order of ldtkldm.py:14-35.  A NaN parameter gives a NaN row."""
    nx, ny, nz = data.shape[:3]
    out = np.empty((len(xs),) + data.shape[3:])
    for j, (x, y, z) in enumerate(zip(xs, ys, zs)):
        if math.isnan(x) or math.isnan(y) or math.isnan(z):
            out[j] = np.nan
            continue
        ix, ax = _grid_cell(x, x0, dx, nx)
        iy, ay = _grid_cell(y, y0, dy, ny)
        iz, az = _grid_cell(z, z0, dz, nz)
        d = data[ix:ix + 2, iy:iy + 2, iz:iz + 2]
        rx, ry, rz = 1.0 - ax, 1.0 - ay, 1.0 - az
        a1, a2, b1, b2 = rx * ry * rz, ax * ry * rz, rx * ay * rz, rx * ry * az
        c1, c2, d1, d2 = ax * ry * az, rx * ay * az, ax * ay * rz, ax * ay * az
        out[j] = (d[0, 0, 0] * a1 + d[1, 0, 0] * a2 + d[0, 1, 0] * b1 + d[0, 0, 1] * b2 +
                  d[1, 0, 1] * c1 + d[0, 1, 1] * c2 + d[1, 1, 0] * d1 + d[1, 1, 1] * d2)
    return out


# --------------------------------------------------------------------------------------
# Orbit: Kepler solve and Taylor-series sky-plane position (orbits_py.py, taylor_z.py)
# --------------------------------------------------------------------------------------

def as_from_rhop(rho: float, period: float) -> float:
    """Scaled semi-major axis from stellar density [g/cc] and period [d]; orbits_py.py:603-618."""
    return (G_NEWTON / (3 * PI)) ** (1 / 3) * ((period * D_S) ** 2 * 1e3 * rho) ** (1 / 3)


def i_from_ba(b: float, a: float) -> float:
    """Inclination from impact parameter and scaled semi-major axis; orbits_py.py:673-688."""
    return math.acos(b / a)


def _pymod(x: float, y: float) -> float:
    """numpy.mod (result has the sign of the divisor)."""
    r = math.fmod(x, y)
    if r != 0.0 and ((r < 0.0) != (y < 0.0)):
        r += y
    elif r == 0.0:
        r = math.copysign(0.0, y)
    return r


def mean_anomaly_offset(e: float, w: float) -> float:
    """orbits_py.py:82-86."""
    m = math.atan2(math.sqrt(1.0 - e ** 2) * math.sin(HALF_PI - w), e + math.cos(HALF_PI - w))
    return m - e * math.sin(m)


def mean_anomaly(t, t0, p, e, w) -> float:
    """orbits_py.py:115-119."""
    offset = mean_anomaly_offset(e, w)
    return _pymod(TWO_PI * (t - (t0 - offset * p / TWO_PI)) / p, TWO_PI)


def ea_newton(t, t0, p, e, w) -> float:
    """Eccentric anomaly by Newton iteration (tol 1e-8, <=1000 steps); orbits_py.py:144-154."""
    ma = mean_anomaly(t, t0, p, e, w)
    ea = ma
    err = 0.05
    k = 0
    while abs(err) > 1e-8 and k < 1000:
        err = ea - e * math.sin(ea) - ma
        ea = ea - err / (1.0 - e * math.cos(ea))
        k += 1
    return ea


def ta_from_ea(ea, e) -> float:
    """orbits_py.py:191-196."""
    sta = math.sqrt(1.0 - e ** 2) * math.sin(ea) / (1.0 - e * math.cos(ea))
    cta = (math.cos(ea) - e) / (1.0 - e * math.cos(ea))
    return math.atan2(sta, cta)


def ta_newton(t, t0, p, e, w) -> float:
    """orbits_py.py:198-200."""
    return ta_from_ea(ea_newton(t, t0, p, e, w), e)


_DT = 2e-2  # taylor_z.py:34


def taylor_coeffs(p, a, i, e, w) -> np.ndarray:
    """Sky-plane Taylor coefficients about mid-transit; taylor_z.py:23-102.

    Seven Newton Kepler solves at t = -3dt..3dt (dt = 0.02 d) and central finite
    difference stencils for velocity, acceleration, jerk and snap.  Returned as
    c[2, 5] with c[:, n] = d^n/dt^n / n!; c[0, 0] (x at mid-transit) is 0 as in
    taylor_z.py:224, c[1, 0] = y3.
    """
    dt = _DT
    ae = a * (1.0 - e ** 2)
    ci = math.cos(i)
    xs = []
    ys = []
    for j in range(-3, 4):
        f = ta_newton(j * dt, 0.0, p, e, w)
        r = ae / (1.0 + e * math.cos(f))
        xs.append(-r * math.cos(w + f))
        ys.append(-r * math.sin(w + f) * ci)
    x0, x1, x2, x3, x4, x5, x6 = xs
    y0, y1, y2, y3, y4, y5, y6 = ys

    a1, b1, c1 = 1 / 60, 9 / 60, 45 / 60
    vx = (a1 * (x6 - x0) + b1 * (x1 - x5) + c1 * (x4 - x2)) / dt
    vy = (a1 * (y6 - y0) + b1 * (y1 - y5) + c1 * (y4 - y2)) / dt
    a2, b2, c2, d2 = 1 / 90, 3 / 20, 3 / 2, 49 / 18
    ax = (a2 * (x0 + x6) - b2 * (x1 + x5) + c2 * (x2 + x4) - d2 * x3) / dt ** 2
    ay = (a2 * (y0 + y6) - b2 * (y1 + y5) + c2 * (y2 + y4) - d2 * y3) / dt ** 2
    a3, b3, c3 = 1 / 8, 1, 13 / 8
    jx = (a3 * (x0 - x6) + b3 * (x5 - x1) + c3 * (x2 - x4)) / dt ** 3
    jy = (a3 * (y0 - y6) + b3 * (y5 - y1) + c3 * (y2 - y4)) / dt ** 3
    a4, b4, c4, d4 = 1 / 6, 2, 13 / 2, 28 / 3
    sx = (-a4 * (x0 + x6) + b4 * (x1 + x5) - c4 * (x2 + x4) + d4 * x3) / dt ** 4
    sy = (-a4 * (y0 + y6) + b4 * (y1 + y5) - c4 * (y2 + y4) + d4 * y3) / dt ** 4

    c = np.zeros((2, 5))
    c[0] = (0.0, vx, 0.5 * ax, jx / 6.0, sx / 24.0)
    c[1] = (y3, vy, 0.5 * ay, jy / 6.0, sy / 24.0)
    return c


def sep(t, c: np.ndarray):
    """Planet-star separation z(t) from Taylor coefficients (sep_c; spec taylor_z.py:229-255)."""
    t = np.asarray(t, float)
    px = (((c[0, 4] * t + c[0, 3]) * t + c[0, 2]) * t + c[0, 1]) * t + c[0, 0]
    py = (((c[1, 4] * t + c[1, 3]) * t + c[1, 2]) * t + c[1, 1]) * t + c[1, 0]
    return np.sqrt(px * px + py * py)


def _z_wrapped(t: float, c: np.ndarray) -> float:
    """z_taylor_s(t, t0=0, p=1, ...): phase-wrapped with a unit period; taylor_z.py:198-226."""
    epoch = math.floor((t - 0.0 + 0.5 * 1.0) / 1.0)
    tt = t - (0.0 + epoch * 1.0)
    return float(sep(tt, c))


def find_contact_point(k: float, point: int, c: np.ndarray) -> float:
    """Contact time by bisection (tol 1e-6 d, <=100 steps); taylor_z.py:298-328."""
    s = -1.0 if point in (1, 2, 12) else 1.0
    if point in (1, 4):
        zt = 1.0 + k
    elif point in (2, 3):
        zt = 1.0 - k
    else:
        zt = 1.0
    vx = c[0, 1]
    t0 = 0.0
    t2 = s * 2.0 / vx
    t1 = 0.5 * t2
    z0 = _z_wrapped(t0, c) - zt
    z1 = _z_wrapped(t1, c) - zt
    i = 0
    while abs(t2 - t0) > 1e-6 and i < 100:
        if z0 * z1 < 0.0:
            t1, t2 = 0.5 * (t0 + t1), t1
            z1 = _z_wrapped(t1, c) - zt
        else:
            t0, t1 = t1, 0.5 * (t1 + t2)
            z0, z1 = z1, _z_wrapped(t1, c) - zt
        i += 1
    return t1


def bounding_box(k: float, c: np.ndarray):
    """First and fourth contact times; taylor_z.py:391-394."""
    return find_contact_point(k, 1, c), find_contact_point(k, 4, c)


# --------------------------------------------------------------------------------------
# RoadRunner model: rr_full (model_full.py:9-100); rr_simple is the npv=nlc=1 case
# --------------------------------------------------------------------------------------

def pv_setup(k, ldp, p, a, i, e, w, nlc, npb, exptimes, weights, dk, kmin, kmax, ze):
    """Per-parameter-vector setup of rr_full; model_full.py:14-70.

    Returns (good[npv], ks[npv,npb], ldm[npv,npb,ng], xyc[npv,2,5], bbs[npv,nlc,2]).
    """
    k = np.asarray(k, float)
    npv = k.shape[0]
    nk, ng = weights.shape[0], weights.shape[1]
    if k.shape[1] > 1 and k.shape[1] != npb:
        raise ValueError('Radius ratios should be given either as an [npv, 1] or [npv, npb] array.')
    ext = np.zeros(nlc)
    ext[:] = exptimes
    ks = np.zeros((npv, npb))
    ks[:, :] = k[:, 0:npb]
    good = np.ones(npv, bool)
    ldm = np.zeros((npv, npb, ng))
    xyc = np.zeros((npv, 2, 5))
    bbs = np.zeros((npv, nlc, 2))
    for ipv in range(npv):
        if math.isnan(a[ipv]) or a[ipv] <= 1.0 or e[ipv] < 0.0 or math.isnan(ldp[ipv, 0, 0]):
            good[ipv] = False
            continue
        k0 = ks[ipv, 0]
        if kmin <= k0 <= kmax:
            ik = int(math.floor((k0 - kmin) / dk))
            ik = min(ik, nk - 2)                      # Q2 (defined)
            ak = (k0 - kmin - ik * dk) / dk
            for ipb in range(npb):
                ldm[ipv, ipb] = (1.0 - ak) * (weights[ik] @ ldp[ipv, ipb]) + ak * (weights[ik + 1] @ ldp[ipv, ipb])
        else:
            _, _, wg = weights_2d(k0, ze, ng)
            for ipb in range(npb):
                ldm[ipv, ipb] = wg @ ldp[ipv, ipb]
        xyc[ipv] = taylor_coeffs(p[ipv], a[ipv], i[ipv], e[ipv], w[ipv])
        bt1, bt4 = bounding_box(k0, xyc[ipv])
        for ilc in range(nlc):
            bbs[ipv, ilc, 0] = bt1 - (0.003 + ext[ilc])
            bbs[ipv, ilc, 1] = bt4 + (0.003 + ext[ilc])
    return good, ks, ldm, xyc, bbs


def rr_full(times, k, t0, p, a, i, e, w, nlc, npb, nep, lcids, pbids, epids, nsamples, exptimes,
            ldp, istar, weights, dk, kmin, kmax, dg, ze):
    """Full RoadRunner model for heterogeneous light curves; model_full.py:9-100.

    Returns flux[npv, npt] (FP64).  The hot loop (model_full.py:75-99) is vectorised
    over the times of one parameter vector; the arithmetic per point is the reference's.
    """
    times = np.asarray(times, float)
    k = np.asarray(k, float)
    t0 = np.asarray(t0, float)
    p, a, i, e, w = (np.atleast_1d(np.asarray(x, float)) for x in (p, a, i, e, w))
    lcids = np.asarray(lcids, np.int64)
    pbids = np.asarray(pbids, np.int64)
    epids = np.asarray(epids, np.int64)
    nsamples = np.atleast_1d(np.asarray(nsamples, np.int64))
    exptimes = np.atleast_1d(np.asarray(exptimes, float))
    npv = k.shape[0]
    npt = times.size
    good, ks, ldm, xyc, bbs = pv_setup(k, ldp, p, a, i, e, w, nlc, npb, exptimes, weights, dk, kmin, kmax, ze)

    ns_lc = np.zeros(nlc, np.int64)
    ns_lc[:] = nsamples
    ex_lc = np.zeros(nlc)
    ex_lc[:] = exptimes
    ipb_t = pbids[lcids]
    iep_t = epids[lcids]

    flux = np.zeros((npv, npt))
    for ipv in range(npv):
        if not good[ipv]:
            flux[ipv] = np.nan
            continue
        tz = t0[ipv, iep_t]
        epoch = np.floor((times - tz + 0.5 * p[ipv]) / p[ipv])
        tc = times - (tz + epoch * p[ipv])
        lo = bbs[ipv, lcids, 0]
        hi = bbs[ipv, lcids, 1]
        inw = (lo <= tc) & (tc <= hi)
        f = np.zeros(npt)
        f[~inw] = 1.0
        for ilc in range(nlc):
            sel = np.nonzero(inw & (lcids == ilc))[0]
            if sel.size == 0:
                continue
            ns = ns_lc[ilc]
            ipb = ipb_t[sel[0]]
            kk = ks[ipv, ipb]
            ist = istar[ipv, ipb]
            acc = np.zeros(sel.size)
            for isample in range(1, ns + 1):
                off = ex_lc[ilc] * ((isample - 0.5) / ns - 0.5)
                z = sep(tc[sel] + off, xyc[ipv])
                ip = interp_ld_mean(z / (1.0 + kk), dg, ldm[ipv, ipb])
                ap = ccia_kite(1.0, kk, z)[0]
                acc += (ist - ip * ap) / ist
            f[sel] = acc / ns
        flux[ipv] = f
    return flux


# --------------------------------------------------------------------------------------
# Model-level drivers mirroring RoadRunnerModel / TSModel (rrmodel.py, tsmodel.py)
# --------------------------------------------------------------------------------------

class OracleRR:
    """RoadRunnerModel restated: construction (rrmodel.py:60-173), set_data
    (transitmodel.py:56-125), evaluate (rrmodel.py:175-238) for the named LD laws."""

    def __init__(self, ldmodel='quadratic', klims=(0.005, 0.5), nk=256, nzin=20, nzlimb=20, zcut=0.7, ng=100):
        self.ldname = ldmodel
        self.klims = klims
        self.nk, self.ng = nk, ng
        self.ze, self.zm = create_z_grid(zcut, nzin, nzlimb)
        self.mu = mu_nodes(self.zm)
        self.dk, self.dg, self.weights = weights_3d(nk, klims[0], klims[1], self.ze, ng)

    def set_data(self, time, lcids=None, pbids=None, nsamples=None, exptimes=None, epids=None):
        self.time = np.asarray(time, float)
        npt = self.time.size
        self.lcids = np.asarray(lcids) if lcids is not None else np.zeros(npt, int)
        self.nlc = np.unique(self.lcids).size
        self.pbids = np.asarray(pbids) if pbids is not None else np.zeros(self.nlc, int)
        self.npb = np.unique(self.pbids).size
        self.epids = np.asarray(epids) if epids is not None else np.zeros(self.nlc, int)
        self.nsamples = np.atleast_1d(nsamples) if nsamples is not None else np.ones(self.nlc, int)
        self.exptimes = np.atleast_1d(exptimes) if exptimes is not None else np.zeros(self.nlc, int)
        self.nep = np.unique(self.epids).size

    def ld(self, ldc):
        return evaluate_ld(self.ldname, self.mu, ldc)

    def evaluate(self, k, ldc, t0, p, a, i, e=0.0, w=0.0, ldp=None, istar=None):
        if ldp is None:
            ldp, istar = self.ld(ldc)
        k = np.asarray(k, float)
        t0 = np.asarray(t0, float)
        if k.ndim == 1:        # Q4 (defined): (npv,) -> (npv, 1)
            k = k[:, None]
        if t0.ndim == 1:
            t0 = t0[:, None]
        k = np.atleast_2d(k)
        t0 = np.atleast_2d(t0)
        p, a, i, e, w = (np.atleast_1d(np.asarray(x, float)) for x in (p, a, i, e, w))
        npv = p.size
        if e.size == 1 and npv > 1:
            e = np.full(npv, e[0])
        if w.size == 1 and npv > 1:
            w = np.full(npv, w[0])
        return rr_full(self.time, k, t0, p, a, i, e, w, self.nlc, self.npb, self.nep, self.lcids, self.pbids,
                       self.epids, self.nsamples, self.exptimes, ldp, istar, self.weights, self.dk,
                       self.klims[0], self.klims[1], self.dg, self.ze)


def tsmodel_serial(times, k, t0, p, a, i, e, w, nsamples, exptimes, ldp, istar, weights, dk, kmin, kmax, ng, dg, ze):
    """TSModel kernel, serial semantics; model_trspec.py:11-93 (Q10: not the parallel variant).

    Returns flux[npv, npb, npt].  ``weights=None`` selects the per-pv exact 2-D
    weights at mean(k) (the default ``precompute_weights=False`` path).  Note the
    reference's per-pv ``kmax = max(k)`` shadows the table bound in the table test.
    """
    times = np.asarray(times, float)
    k = np.asarray(k, float)
    if k.ndim != 2:
        raise ValueError(" The radius ratios must be given as a 2D array with shape (npv, npb)")
    if ldp.ndim != 3:
        raise ValueError("The limb darkening profiles must be given as a 3D array with shape (npv, npb, nmu)")
    if k.shape[1] != ldp.shape[1]:
        raise ValueError("The transmission spectrum transit model requires that the number or radius ratios "
                         "and the number of passbands match.")
    t0, p, a, i, e, w = (np.atleast_1d(np.asarray(x, float)) for x in (t0, p, a, i, e, w))
    nsamples = np.atleast_1d(nsamples)
    exptimes = np.atleast_1d(exptimes)
    npt = times.size
    npv, npb = k.shape
    if weights is not None:
        ng = weights.shape[1]
        nk = weights.shape[0]
    flux = np.zeros((npv, npb, npt))
    ns = int(nsamples[0])
    ex0 = float(exptimes[0])
    for ipv in range(npv):
        if math.isnan(a[ipv]) or a[ipv] <= 1.0 or e[ipv] < 0.0:
            flux[ipv] = np.nan
            continue
        kmean = np.cumsum(k[ipv])[-1] / npb
        kmax_pv = np.max(k[ipv])
        afac = k[ipv] ** 2 / kmean ** 2
        # The reference tests kmin <= kmean <= max(k) (its per-pv kmax shadows the table
        # bound), which holds for any kmean >= kmin and then reads past the table when
        # kmean > kmax.  Defined here: the table serves kmean in [kmin, kmax] (table
        # bounds); outside it the exact 2-D weights at kmean are used, as RR does.
        if weights is not None and kmin <= kmean <= kmax:
            ik = int(math.floor((kmean - kmin) / dk))
            ik = min(ik, nk - 2)
            ak = (kmean - kmin - ik * dk) / dk
            ldm = (1.0 - ak) * (ldp[ipv] @ weights[ik].T) + ak * (ldp[ipv] @ weights[ik + 1].T)
            dgv = dg
        else:
            _, dgv, wg = weights_2d(kmean, ze, ng)
            ldm = ldp[ipv] @ wg.T                      # [npb, ng]
        xyc = taylor_coeffs(p[ipv], a[ipv], i[ipv], e[ipv], w[ipv])
        bt1, bt4 = bounding_box(kmean, xyc)
        bt1 -= 0.0015 + ex0
        bt4 += 0.0015 + ex0
        epoch = np.floor((times - t0[ipv] + 0.5 * p[ipv]) / p[ipv])
        tc = times - (t0[ipv] + epoch * p[ipv])
        inw = (bt1 <= tc) & (tc <= bt4)
        f = np.zeros((npb, npt))
        f[:, ~inw] = 1.0
        sel = np.nonzero(inw)[0]
        if sel.size:
            acc = np.zeros((npb, sel.size))
            for isample in range(1, ns + 1):
                off = ex0 * ((isample - 0.5) / ns - 0.5)
                z = sep(tc[sel] + off, xyc)
                ap0, kap = ccia_kite(1.0, kmean, z)
                dadk = 2.0 * kmean * kap
                g = z / (1.0 + kmean)
                inside = z <= 1.0 - kmax_pv
                for ipb in range(npb):
                    ip = interp_ld_mean(g, dgv, ldm[ipb])
                    ist = istar[ipv, ipb]
                    area = np.where(inside, ap0 * afac[ipb], ap0 + (k[ipv, ipb] - kmean) * dadk)
                    acc[ipb] += (ist - ip * area) / ist
            f[:, sel] = acc / ns
        flux[ipv] = f
    return flux


# --------------------------------------------------------------------------------------
# Mandel & Agol (2002) quadratic model: the second flux oracle (numba/ma_quadratic_nb.py)
# --------------------------------------------------------------------------------------

def _ellpicb(n, k):
    """Complete elliptic integral of the third kind (Bulirsch); ma_quadratic_nb.py:55-91."""
    kc = math.sqrt(1.0 - k * k)
    e = kc
    p = math.sqrt(n + 1.0)
    m0 = 1.0
    c = 1.0
    d = 1.0 / p
    for _ in range(1000):
        f = c
        c = d / p + c
        g = e / p
        d = 2.0 * (f * g + d)
        p = g + p
        g = m0
        m0 = kc + m0
        if abs(1.0 - kc / g) > 1e-8:
            kc = 2.0 * math.sqrt(e)
            e = kc * m0
        else:
            return HALF_PI * (c * m0 + d) / (m0 * (m0 + p))
    return 0.0


def _ellec(k):
    """Hastings polynomial approximation of E(k); ma_quadratic_nb.py:93-108."""
    m1 = 1.0 - k * k
    ee1 = 1.0 + m1 * (0.443251414630 + m1 * (0.062606012200 + m1 * (0.047573835460 + m1 * 0.017365064510)))
    ee2 = m1 * (0.249983683100 + m1 * (0.092001800370 + m1 * (0.040696975260 + m1 * 0.005264496390))) * math.log(1.0 / m1)
    return ee1 + ee2


def _ellk(k):
    """Hastings polynomial approximation of K(k); ma_quadratic_nb.py:110-126."""
    m1 = 1.0 - k * k
    ek1 = 1.386294361120 + m1 * (0.096663442590 + m1 * (0.035900923830 + m1 * (0.037425637130 + m1 * 0.014511962120)))
    ek2 = (0.50 + m1 * (0.124985935970 + m1 * (0.068802485760 + m1 * (0.033283553460 + m1 * 0.004417870120)))) * math.log(m1)
    return ek1 - ek2


def eval_quad_z(z: float, k: float, u) -> float:
    """Mandel-Agol quadratic-LD flux at separation z; ma_quadratic_nb.py:277-389."""
    inv_pi = 1.0 / PI
    if abs(k - 0.5) < 1.0e-4:
        k = 0.5
    k2 = k ** 2
    omega = 1.0 - u[0] / 3.0 - u[1] / 6.0
    if abs(z - k) < 1e-6:
        z += 1e-6
    if z > 1.0 + k or math.copysign(1, z) < 0.0:
        return 1.0
    elif k >= 1.0 and z <= k - 1.0:
        return 0.0
    z2 = z ** 2
    x1 = (k - z) ** 2
    x2 = (k + z) ** 2
    x3 = k ** 2 - z ** 2
    kap0 = kap1 = 0.0
    le = 0.0
    if z <= 1.0 - k:
        le = k2
    elif abs(1.0 - k) <= z <= 1.0 + k:
        kap1 = math.acos(min((1.0 - k2 + z2) / (2.0 * z), 1.0))
        kap0 = math.acos(min((k2 + z2 - 1.0) / (2.0 * k * z), 1.0))
        le = k2 * kap0 + kap1
        le = (le - 0.5 * math.sqrt(max(4.0 * z2 - (1.0 + z2 - k2) ** 2, 0.0))) * inv_pi
    edge_at_origin = abs(z - k) < 1.e-4 * (z + k)
    full_transit = k <= 1.0 and z < (1.0 - k)
    if edge_at_origin:
        if k == 0.5:
            ld = 1.0 / 3.0 - 4.0 * inv_pi / 9.0
            ed = 3.0 / 32.0
        elif z > 0.5:
            q = 0.50 / k
            kk, ek = _ellk(q), _ellec(q)
            ld = (1.0 / 3.0 + 16.0 * k / 9.0 * inv_pi * (2.0 * k2 - 1.0) * ek
                  - (32.0 * k ** 4 - 20.0 * k2 + 3.0) / 9.0 * inv_pi / k * kk)
            ed = 1.0 / 2.0 * inv_pi * (kap1 + k2 * (k2 + 2.0 * z2) * kap0
                                       - (1.0 + 5.0 * k2 + z2) / 4.0 * math.sqrt((1.0 - x1) * (x2 - 1.0)))
        else:
            q = 2.0 * k
            kk, ek = _ellk(q), _ellec(q)
            ld = 1.0 / 3.0 + 2.0 / 9.0 * inv_pi * (4.0 * (2.0 * k2 - 1.0) * ek + (1.0 - 4.0 * k2) * kk)
            ed = k2 / 2.0 * (k2 + 2.0 * z2)
    else:
        if full_transit:
            q = math.sqrt((x2 - x1) / (1.0 - x1))
            kk, ek = _ellk(q), _ellec(q)
            n = x2 / x1 - 1.0
            pk = _ellpicb(n, q)
            ld = 2.0 / 9.0 * inv_pi / math.sqrt(1.0 - x1) * ((1.0 - 5.0 * z2 + k2 + x3 * x3) * kk
                                                              + (1.0 - x1) * (z2 + 7.0 * k2 - 4.0) * ek
                                                              - 3.0 * x3 / x1 * pk)
            if z < k:
                ld = ld + 2.0 / 3.0
            if abs(k + z - 1.0) < 1.e-4:
                ld = (2.0 / 3.0 * inv_pi * math.acos(1.0 - 2.0 * k)
                      - 4.0 / 9.0 * inv_pi * math.sqrt(k * (1.0 - k)) * (3.0 + 2.0 * k - 8.0 * k2))
            ed = k2 / 2.0 * (k2 + 2.0 * z2)
        else:
            q = math.sqrt((1.0 - (k - z) ** 2) / 4.0 / z / k)
            kk, ek = _ellk(q), _ellec(q)
            n = 1.0 / x1 - 1.0
            pk = _ellpicb(n, q)
            ld = 1.0 / 9.0 * inv_pi / math.sqrt(k * z) * (
                ((1.0 - x2) * (2.0 * x2 + x1 - 3.0) - 3.0 * x3 * (x2 - 2.0)) * kk
                + 4.0 * k * z * (z2 + 7.0 * k2 - 4.0) * ek - 3.0 * x3 / x1 * pk)
            if z < k:
                ld = ld + 2.0 / 3.0
            ed = 1.0 / 2.0 * inv_pi * (kap1 + k2 * (k2 + 2.0 * z2) * kap0
                                       - (1.0 + 5.0 * k2 + z2) / 4.0 * math.sqrt((1.0 - x1) * (x2 - 1.0)))
    return 1.0 - ((1.0 - u[0] - 2.0 * u[1]) * le + (u[0] + 2.0 * u[1]) * ld + u[1] * ed) / omega


def quadratic_model(times, k, t0, p, a, i, e, w, ldc, lcids, pbids, epids, nsamples, exptimes):
    """Mandel-Agol light curves of a parameter population; quadratic_model_v,
    ma_quadratic_nb.py:532-595 (the non-interpolated branch).  k [npv, 1|npb],
    t0 [npv, nep], ldc [npv, 2*npb]; flux [npv, npt].  A NaN k or a gives an inf
    row, NaN coefficients inf points; the window is the contact times +- 0.025 d."""
    times = np.asarray(times, float)
    k, t0, ldc = np.atleast_2d(k), np.atleast_2d(t0), np.atleast_2d(ldc)
    npv, npt = k.shape[0], times.size
    flux = np.zeros((npv, npt))
    for ipv in range(npv):
        if np.isnan(k[ipv]).any() or math.isnan(a[ipv]):
            flux[ipv] = np.inf
            continue
        c = taylor_coeffs(p[ipv], a[ipv], i[ipv], e[ipv], w[ipv])
        bt1, bt4 = bounding_box(k[ipv, 0], c)
        bt1 -= 0.025
        bt4 += 0.025
        for j in range(npt):
            ilc = lcids[j]
            iep = epids[ilc]
            epoch = math.floor((times[j] - t0[ipv, iep] + 0.5 * p[ipv]) / p[ipv])
            tc = times[j] - (t0[ipv, iep] + epoch * p[ipv])
            if not (bt1 <= tc <= bt4):
                flux[ipv, j] = 1.0
                continue
            ipb = pbids[ilc]
            kk = k[ipv, 0] if k.shape[1] == 1 else k[ipv, ipb]
            u = ldc[ipv, 2 * ipb:2 * (ipb + 1)]
            if np.isnan(u).any():
                flux[ipv, j] = np.inf
                continue
            acc = 0.0
            ns = nsamples[ilc]
            for s in range(1, ns + 1):
                z = sep(tc + exptimes[ilc] * ((s - 0.5) / ns - 0.5), c)
                acc += 1.0 if z > 1.0 + kk else eval_quad_z(z, kk, u)
            flux[ipv, j] = acc / ns
    return flux
