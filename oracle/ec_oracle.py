"""FP64 numpy restatement of PyTransit's RoadRunner eclipse, eclipse-spectroscopy
and oblate-planet variants.

TEST INFRASTRUCTURE ONLY, like ``rr_oracle``: only ``tests/`` may import it; the
product path (``pytransit_amd``) never calls it.

Functions and the reference file:line each restates (paths relative to the
PyTransit 2.8.1 checkout):

* ``eclipse_phase``           orbits/orbits_py.py:544-555
* ``taylor_coeffs_eclipse``   orbits/taylor_z.py:106-185 (vajs_from_paiew_eclipse)
* ``light_travel_time``       meepmeep ``eclipse_light_travel_time`` (external; see below)
* ``bounding_box_eclipse``    orbits/taylor_z.py:298-328,391-394, bracket by |vx| (see below)
* ``eclipse_model``           models/roadrunner/model_eclipse.py:11-81
* ``esmodel``                 models/roadrunner/model_ecspec.py:13-65
* ``create_ellipse``          models/roadrunner/ecintersection.py:93-134
* ``ellipse_circle_area``     models/roadrunner/ecintersection.py:137-207
* ``opmodel``                 models/roadrunner/opmodel_full.py:11-134

The eclipse kernels call three meepmeep functions that are not in the reference
checkout (``eclipse_time_offset``, ``solve2d`` about the eclipse, and
``eclipse_light_travel_time``).  They are restated from PyTransit's own in-repo
methods and from the geometry:

* the eclipse time offset is ``eclipse_phase`` (the exact conjunction phase of
  orbits_py.py:544-555), and the Taylor series about mid-eclipse is the in-repo
  stencil ``vajs_from_paiew_eclipse`` (taylor_z.py:106-185);
* the light travel time is the extra light path between the two conjunctions,
  (r_transit + r_eclipse) sin(i) R_star / c, with r the star-planet distance at
  each conjunction (R_sun = 6.957e8 m, c = 299 792 458 m/s);
* the contact times use the in-repo bisection (taylor_z.py:298-328) with the
  bracket taken on the side given by the sign of the contact (before or after
  mid-eclipse) rather than by the sign of vx: behind the star the planet moves in
  -x, and the in-repo bracket s*2/vx would then place first contact after mid-
  eclipse and leave an empty window, which the reference's own eclipse test
  (tests/test_roadrunner_eclipse.py:57-69, a dip at p/2) rules out.  For a
  transit (vx > 0) the two brackets are the same.

These three are pinned here by the mathematics (closed forms and brute force on
the oracle's pinned Kepler solver), not against meepmeep: parity with meepmeep at
this boundary is unpinned, as it is for the transit path's ``solve2d``.
"""
import math

import numpy as np

from .rr_oracle import (PI, _DT, bounding_box, ccia_kite, interp_ld_mean, sep, ta_newton, taylor_coeffs,
                        weights_2d)

R_SUN_M = 6.957e8          # IAU 2015 nominal solar radius [m]
C_LIGHT = 299792458.0      # [m/s]
D_S = 86400.0              # [s/d]


# --------------------------------------------------------------------------------------
# Orbit about the secondary eclipse
# --------------------------------------------------------------------------------------

def eclipse_phase(p, i, e, w):
    """Time from mid-transit to mid-eclipse (the conjunctions); orbits_py.py:544-555."""
    etr = math.atan2(math.sqrt(1. - e ** 2) * math.sin(0.5 * PI - w), e + math.cos(0.5 * PI - w))
    eec = math.atan2(math.sqrt(1. - e ** 2) * math.sin(0.5 * PI + PI - w), e + math.cos(0.5 * PI + PI - w))
    mtr = etr - e * math.sin(etr)
    mec = eec - e * math.sin(eec)
    phase = (mec - mtr) * p / (2 * PI)
    return phase if phase > 0. else p + phase


def _stencil_xy(t, p, a, ci, e, w):
    f = ta_newton(t, 0.0, p, e, w)
    r = a * (1.0 - e ** 2) / (1.0 + e * math.cos(f))
    return -r * math.cos(w + f), -r * math.sin(w + f) * ci


def taylor_coeffs_eclipse(p, a, i, e, w):
    """(te, c[2, 5]): the Taylor series of the sky position about mid-eclipse;
    taylor_z.py:106-185.  Seven Newton Kepler solves at te + j dt (j = -3..3,
    dt = 0.02 d) and the same central stencils as the transit series; the x
    constant term is 0 (x vanishes at conjunction)."""
    te = eclipse_phase(p, i, e, w)
    ci = math.cos(i)
    xs, ys = zip(*(_stencil_xy(te + j * _DT, p, a, ci, e, w) for j in range(-3, 4)))
    return te, _coeffs_from_stencil(xs, ys)


def _coeffs_from_stencil(xs, ys):
    # the stencil of taylor_z.py:78-102, as rr_oracle.taylor_coeffs applies it
    dt = _DT
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


def light_travel_time(p, a, i, e, w, rstar):
    """Extra light travel time of the eclipse relative to the transit [d]: the
    planet's line-of-sight distance behind the star at eclipse plus its distance
    in front at transit, (r_tr + r_ec) sin i, in stellar radii of ``rstar`` R_sun."""
    r_tr = a * (1.0 - e ** 2) / (1.0 + e * math.sin(w))
    r_ec = a * (1.0 - e ** 2) / (1.0 - e * math.sin(w))
    return (r_tr + r_ec) * math.sin(i) * rstar * R_SUN_M / C_LIGHT / D_S


def _z_unit_wrapped(t, c):
    epoch = math.floor((t - 0.0 + 0.5 * 1.0) / 1.0)
    return float(sep(t - (0.0 + epoch * 1.0), c))


def _contact(k, point, c):
    # taylor_z.py:298-328 with the bracket on the contact's side of mid-eclipse
    s = -1.0 if point in (1, 2, 12) else 1.0
    zt = 1.0 + k if point in (1, 4) else (1.0 - k if point in (2, 3) else 1.0)
    t0 = 0.0
    t2 = s * 2.0 / abs(c[0, 1])
    t1 = 0.5 * t2
    z0 = _z_unit_wrapped(t0, c) - zt
    z1 = _z_unit_wrapped(t1, c) - zt
    i = 0
    while abs(t2 - t0) > 1e-6 and i < 100:
        if z0 * z1 < 0.0:
            t1, t2 = 0.5 * (t0 + t1), t1
            z1 = _z_unit_wrapped(t1, c) - zt
        else:
            t0, t1 = t1, 0.5 * (t1 + t2)
            z0, z1 = z1, _z_unit_wrapped(t1, c) - zt
        i += 1
    return t1


def bounding_box_eclipse(k, c):
    """First and fourth contact times about mid-eclipse (see the module docstring)."""
    return _contact(k, 1, c), _contact(k, 4, c)


def _eclipse_setup(k, p, a, i, e, w, rstar):
    """Per-pv (good, shift, c, bt1, bt4) of model_eclipse.py:32-53 / model_ecspec.py:29-48."""
    if math.isnan(a) or a <= 1.0 or e < 0.0:
        return False, 0.0, None, 0.0, 0.0
    te, c = taylor_coeffs_eclipse(p, a, i, e, w)
    shift = te + light_travel_time(p, a, i, e, w, rstar)
    bt1, bt4 = bounding_box_eclipse(k, c)
    return True, shift, c, bt1, bt4


# --------------------------------------------------------------------------------------
# Eclipse and eclipse-spectroscopy models
# --------------------------------------------------------------------------------------

def eclipse_model(times, k, t0, p, a, i, e, w, rstar, nlc, lcids, epids, nsamples, exptimes):
    """Visible planet disk area pi k^2 - A(1, k, z) over the secondary eclipse;
    model_eclipse.py:11-81.  k [npv], t0 [npv, nep]; returns [npv, npt]."""
    times = np.asarray(times, float)
    k = np.asarray(k, float)
    if k.ndim != 1:
        raise ValueError(" The radius ratio must be given as an 1D array with shape (npv)")
    t0 = np.asarray(t0, float)
    p, a, i, e, w = (np.asarray(x, float) for x in (p, a, i, e, w))
    lcids = np.asarray(lcids, np.int64)
    epids = np.asarray(epids, np.int64)
    ns_lc = np.zeros(nlc, np.int64)
    ns_lc[:] = nsamples
    ex_lc = np.zeros(nlc)
    ex_lc[:] = exptimes
    npv, npt = k.size, times.size
    flux = np.zeros((npv, npt))
    for ipv in range(npv):
        good, shift, c, bt1, bt4 = _eclipse_setup(k[ipv], p[ipv], a[ipv], i[ipv], e[ipv], w[ipv], rstar)
        if not good:
            flux[ipv] = np.nan
            continue
        k2 = PI * k[ipv] ** 2
        te = t0[ipv, epids[lcids]] + shift
        epoch = np.floor((times - te + 0.5 * p[ipv]) / p[ipv])
        tc = times - (te + epoch * p[ipv])
        lo = bt1 - (0.003 + ex_lc[lcids])
        hi = bt4 + (0.003 + ex_lc[lcids])
        inw = (lo <= tc) & (tc <= hi)
        f = np.full(npt, k2)
        for ilc in range(nlc):
            sel = np.nonzero(inw & (lcids == ilc))[0]
            if sel.size == 0:
                continue
            ns = ns_lc[ilc]
            acc = np.zeros(sel.size)
            for isample in range(1, ns + 1):
                off = ex_lc[ilc] * ((isample - 0.5) / ns - 0.5)
                z = sep(tc[sel] + off, c)
                acc += k2 - ccia_kite(1.0, k[ipv], z)[0]
            f[sel] = acc / ns
        flux[ipv] = f
    return flux


def esmodel(times, k, t0, p, a, i, e, w, rstar, fratio, nsamples, exptime):
    """Eclipse spectroscopy: flux[npv, npb, npt] = 1 - (f A / pi) / (1 + f k^2)
    averaged over the exposure, 1 outside the window; model_ecspec.py:13-65."""
    times = np.asarray(times, float)
    k = np.asarray(k, float)
    fratio = np.asarray(fratio, float)
    if k.ndim != 1:
        raise ValueError(" The radius ratio must be given as an 1D array with shape (npv)")
    if fratio.ndim != 2:
        raise ValueError(" The flux ratio must be given as a 2D array with shape (npv, npb)")
    t0, p, a, i, e, w, rstar = (np.asarray(x, float) for x in (t0, p, a, i, e, w, rstar))
    npt = times.size
    npv, npb = fratio.shape
    flux = np.zeros((npv, npb, npt))
    for ipv in range(npv):
        good, shift, c, bt1, bt4 = _eclipse_setup(k[ipv], p[ipv], a[ipv], i[ipv], e[ipv], w[ipv], rstar[ipv])
        if not good:
            flux[ipv] = np.nan
            continue
        te = t0[ipv] + shift
        lo, hi = bt1 - (0.0015 + exptime), bt4 + (0.0015 + exptime)
        epoch = np.floor((times - te + 0.5 * p[ipv]) / p[ipv])
        tc = times - (te + epoch * p[ipv])
        inw = (lo <= tc) & (tc <= hi)
        f = np.ones((npb, npt))
        sel = np.nonzero(inw)[0]
        if sel.size:
            acc = np.zeros((npb, sel.size))
            fr = fratio[ipv][:, None]
            for isample in range(1, nsamples + 1):
                off = exptime * ((isample - 0.5) / nsamples - 0.5)
                z = sep(tc[sel] + off, c)
                ar = ccia_kite(1.0, k[ipv], z)[0][None, :]
                acc += 1.0 - (fr * ar / PI) / (1.0 + fr * k[ipv] ** 2)
            f[:, sel] = acc / nsamples
        flux[ipv] = f
    return flux


# --------------------------------------------------------------------------------------
# Oblate planet: a rotated ellipse crossing the stellar disk
# --------------------------------------------------------------------------------------

def create_ellipse(ny, k, f, a):
    """Scanlines of the planet ellipse; ecintersection.py:93-134.

    Semi-axes k (along x before rotation) and k (1 - f), rotated by ``a``.  Returns
    ys[ny] (row centres, k (j dy - 1) with dy = 2 / (ny + 1)) and xs[ny, 2] (left and
    right ends of each row, NaN where the row misses the ellipse)."""
    dy = 2 / (ny + 1)
    ys = np.arange(1, 1 + ny) * dy - 1
    xs = np.full((ny, 2), np.nan)
    ca, sa = math.cos(a), math.sin(a)
    b = 1.0 - f
    b2, ca2, sa2 = b * b, ca * ca, sa * sa
    for j in range(ny):
        y = ys[j]
        y2 = y * y
        d = b2 * ca2 + ca2 * ca2 * (-y2) - 2 * ca2 * sa2 * y2 - sa2 * sa2 * y2 + sa2
        if d >= 0:
            d = math.sqrt(d) / b
            u = (ca * sa * y) / b2 - ca * sa * y
            v = sa2 / b2 + ca2
            xs[j, 0] = (-d + u) / v
            xs[j, 1] = (d + u) / v
    return k * ys, k * xs


def ellipse_circle_area(cx, cy, z, k, f, xs, ys):
    """Area of the planet ellipse (centre (cx, cy), rows xs/ys about its centre)
    inside the unit stellar disk by scanline sums; ecintersection.py:137-207, with
    its three branches by the ellipse centre's x."""
    if z <= 1.0 - k:
        return PI * k * (1.0 - f) * k
    if z >= 1.0 + k:
        return 0.0
    ny = ys.size
    dy = ys[1] - ys[0]
    acc = 0.0
    with np.errstate(invalid='ignore'):
        if cx < -k:
            for j in range(ny):
                if math.isfinite(xs[j, 0]):
                    xstar = -np.sqrt(1.0 - (ys[j] + cy) ** 2) - cx
                    if xstar <= xs[j, 0]:
                        acc += xs[j, 1] - xs[j, 0]
                    elif xstar <= xs[j, 1]:
                        acc += xs[j, 1] - xstar
        elif cx > k:
            for j in range(ny):
                if math.isfinite(xs[j, 0]):
                    xstar = np.sqrt(1.0 - (ys[j] + cy) ** 2) - cx
                    if xstar >= xs[j, 1]:
                        acc += xs[j, 1] - xs[j, 0]
                    elif xstar >= xs[j, 0]:
                        acc += xstar - xs[j, 0]
        else:
            for j in range(ny):
                if math.isfinite(xs[j, 0]):
                    if abs(ys[j] + cy) <= 1.0:
                        xstar = math.sqrt(1.0 - (ys[j] + cy) ** 2)
                        xst1 = -xstar - cx
                        xst2 = xstar - cx
                        if xst1 <= xs[j, 0]:
                            acc += min(xst2, xs[j, 1]) - xs[j, 0]
                        elif xst1 > xs[j, 0]:
                            acc += min(xst2, xs[j, 1]) - xst1
    return acc * dy


def opmodel(times, k, f, alpha, t0, p, a, i, e, w, nlc, npb, nep, npl, lcids, pbids, epids, nsamples, exptimes,
            ldp, istar, weights, dk, kmin, kmax, dg, ze):
    """Oblate-planet RoadRunner model; opmodel_full.py:40-134 (op_full_serial).

    k [npv, 1|npb], f, alpha [npv] (flattening and sky-plane rotation), t0
    [npv, nep]; returns flux [npv, npt].  The limb-darkening means are rr_full's
    (the table lerp at k[:, 0], Q2's clamp defined as in rr_oracle); the window
    margin is 0.0015 + exptime; the ellipse rows are built at k[:, 0] and the area
    uses each passband's k for its full/none branches."""
    times = np.asarray(times, float)
    k = np.atleast_2d(np.asarray(k, float))
    t0 = np.atleast_2d(np.asarray(t0, float))
    f, alpha, p, a, i, e, w = (np.atleast_1d(np.asarray(x, float)) for x in (f, alpha, p, a, i, e, w))
    lcids = np.asarray(lcids, np.int64)
    pbids = np.asarray(pbids, np.int64)
    epids = np.asarray(epids, np.int64)
    npv, npt = k.shape[0], times.size
    nk, ng = weights.shape[0], weights.shape[1]
    if k.shape[1] > 1 and k.shape[1] != npb:
        raise ValueError('Radius ratios should be given either as an [npv, 1] or [npv, npb] array.')
    ns_lc = np.zeros(nlc, np.int64)
    ns_lc[:] = nsamples
    ex_lc = np.zeros(nlc)
    ex_lc[:] = exptimes
    ks = np.zeros((npv, npb))
    ks[:, :] = k[:, 0:npb]
    flux = np.zeros((npv, npt))
    for ipv in range(npv):
        if math.isnan(a[ipv]) or a[ipv] <= 1.0 or e[ipv] < 0.0 or math.isnan(ldp[ipv, 0, 0]):
            flux[ipv] = np.nan
            continue
        k0 = ks[ipv, 0]
        ldm = np.zeros((npb, ng))
        if kmin <= k0 <= kmax:
            ik = min(int(math.floor((k0 - kmin) / dk)), nk - 2)
            ak = (k0 - kmin - ik * dk) / dk
            for ipb in range(npb):
                ldm[ipb] = (1.0 - ak) * (weights[ik] @ ldp[ipv, ipb]) + ak * (weights[ik + 1] @ ldp[ipv, ipb])
        else:
            _, _, wg = weights_2d(k0, ze, ng)
            for ipb in range(npb):
                ldm[ipb] = wg @ ldp[ipv, ipb]
        c = taylor_coeffs(p[ipv], a[ipv], i[ipv], e[ipv], w[ipv])
        bt1, bt4 = bounding_box(k0, c)
        ey, ex = create_ellipse(npl, k0, f[ipv], alpha[ipv])
        for ipt in range(npt):
            ilc = lcids[ipt]
            ipb = pbids[ilc]
            tz = t0[ipv, epids[ilc]]
            epoch = math.floor((times[ipt] - tz + 0.5 * p[ipv]) / p[ipv])
            tc = times[ipt] - (tz + epoch * p[ipv])
            if not (bt1 - (0.0015 + ex_lc[ilc]) <= tc <= bt4 + (0.0015 + ex_lc[ilc])):
                flux[ipv, ipt] = 1.0
                continue
            ns = ns_lc[ilc]
            acc = 0.0
            for isample in range(1, ns + 1):
                tt = tc + ex_lc[ilc] * ((isample - 0.5) / ns - 0.5)
                cx = (((c[0, 4] * tt + c[0, 3]) * tt + c[0, 2]) * tt + c[0, 1]) * tt + c[0, 0]
                cy = (((c[1, 4] * tt + c[1, 3]) * tt + c[1, 2]) * tt + c[1, 1]) * tt + c[1, 0]
                z = math.sqrt(cx * cx + cy * cy)
                ip = float(interp_ld_mean(z / (1.0 + ks[ipv, ipb]), dg, ldm[ipb]))
                ap = ellipse_circle_area(cx, cy, z, ks[ipv, ipb], f[ipv], ex, ey)
                acc += (istar[ipv, ipb] - ip * ap) / istar[ipv, ipb]
            flux[ipv, ipt] = acc / ns
    return flux
