"""Pins of the eclipse, eclipse-spectroscopy and oblate-planet restatements
(oracle/ec_oracle.py), each against something the restatement does not itself
compute: the reference's own eclipse test, closed forms, the pinned Kepler
solver and lens area of rr_oracle, and brute-force areas on a grid.

Parity with meepmeep's eclipse_time_offset / solve2d / eclipse_light_travel_time
is unpinned (meepmeep is not in the reference checkout and not installed); the
restatement follows PyTransit's in-repo methods and the geometry, pinned here.
"""
import math

import numpy as np
import pytest

from oracle import ec_oracle as eo
from oracle import rr_oracle as ro


def reference_eclipse_inputs(k=0.1, a=8.0, inc=np.pi / 2, p=2.0, t0=0.0, e=0.0, w=0.0, rstar=1.0, npt=5000,
                             t_start=-0.5, t_end=3.5):
    # tests/test_roadrunner_eclipse.py:16-41 (make_inputs)
    times = np.linspace(t_start, t_end, npt)
    return (times, np.array([k]), np.full((1, 1), t0), np.array([p]), np.array([a]), np.array([inc]),
            np.array([e]), np.array([w]), float(rstar), 1, np.zeros(npt, np.int64), np.zeros(1, np.int64),
            np.ones(1, np.int64), np.zeros(1))


def test_reference_eclipse_test_properties():
    """tests/test_roadrunner_eclipse.py:50-69: shape, pi k^2 out of eclipse, a dip
    at p/2 and a total eclipse reaching ~0."""
    args = reference_eclipse_inputs()
    times, k = args[0], args[1][0]
    flux = eo.eclipse_model(*args)
    assert flux.shape == (1, times.size)
    f = flux[0]
    base = math.pi * k ** 2
    assert f.max() == pytest.approx(base, abs=1e-6)
    assert f[np.argmin(np.abs(times - 1.0))] < base - 1e-4
    assert f.min() < 1e-3
    # the primary transit at t = 0 is not an eclipse: the area stays pi k^2 there
    assert np.all(f[np.abs(times) < 0.2] == base)


@pytest.mark.parametrize('p,e,w', [(2.0, 0.0, 0.0), (3.3, 0.2, 0.7), (7.1, 0.45, -2.1), (1.4, 0.05, 3.0)])
def test_eclipse_phase_is_the_far_conjunction(p, e, w):
    """At te the true anomaly of the pinned Kepler solver puts the planet at
    w + f = 3 pi / 2: on the sky axis x = 0, behind the star (y_los > 0)."""
    te = eo.eclipse_phase(p, 1.5, e, w)
    assert 0.0 < te < p
    if e == 0.0:
        assert te == pytest.approx(0.5 * p, rel=1e-15)
    f = ro.ta_newton(te, 0.0, p, e, w)
    ang = (w + f) % (2 * math.pi)
    assert ang == pytest.approx(1.5 * math.pi, abs=1e-7)
    # and the transit is at w + f = pi / 2 at t = 0
    assert (w + ro.ta_newton(0.0, 0.0, p, e, w)) % (2 * math.pi) == pytest.approx(0.5 * math.pi, abs=1e-7)


def test_eclipse_taylor_series_follows_kepler():
    """The series about mid-eclipse follows the exact orbit about te, and for e = 0
    it is the transit series with vx and the y row mirrored."""
    p, a, inc, e, w = 2.5, 7.0, 1.50, 0.3, 1.1
    te, c = eo.taylor_coeffs_eclipse(p, a, inc, e, w)
    t = np.linspace(-0.05, 0.05, 21)
    exact = []
    for tt in t:
        f = ro.ta_newton(te + tt, 0.0, p, e, w)
        r = a * (1 - e ** 2) / (1 + e * math.cos(f))
        exact.append(math.hypot(-r * math.cos(w + f), -r * math.sin(w + f) * math.cos(inc)))
    err = np.abs(ro.sep(t, c) - np.array(exact))
    assert err.max() < 1e-4 and err[np.abs(t) <= 0.02].max() < 1e-6
    _, ce = eo.taylor_coeffs_eclipse(3.0, 9.0, 1.52, 0.0, 0.0)
    ct = ro.taylor_coeffs(3.0, 9.0, 1.52, 0.0, 0.0)
    np.testing.assert_allclose(ce[0, 1], -ct[0, 1], rtol=1e-9)   # moving in -x behind the star
    np.testing.assert_allclose(ce[1, 0], -ct[1, 0], rtol=1e-9)   # y mirrored
    assert ce[0, 1] < 0.0


def test_eclipse_contact_times_closed_form():
    """e = 0: the eclipse lasts T14 = (p/pi) asin(sqrt((1+k)^2 - b^2)/(a sin i)),
    symmetric about mid-eclipse, first contact before it."""
    for p, a, b, k in ((1.0, 3.0, 0.0, 0.1), (4.2, 11.0, 0.5, 0.08)):
        inc = math.acos(b / a)
        _, c = eo.taylor_coeffs_eclipse(p, a, inc, 0.0, 0.0)
        t1, t4 = eo.bounding_box_eclipse(k, c)
        t14 = p / math.pi * math.asin(math.sqrt((1 + k) ** 2 - b ** 2) / (a * math.sin(inc)))
        assert t1 < 0.0 < t4
        assert t4 - t1 == pytest.approx(t14, rel=2e-4)
        assert float(ro.sep(t4, c)) == pytest.approx(1 + k, abs=2e-5)


def test_light_travel_time_is_the_line_of_sight_path():
    """The delay equals the line-of-sight separation of the planet's positions at
    the two conjunctions, from the pinned Kepler solver: for e = 0, 2 a sin i R*/c."""
    au_rsun = 215.032
    assert eo.light_travel_time(365.25, au_rsun, math.pi / 2, 0.0, 0.0, 1.0) * 86400 == pytest.approx(
        2 * 499.005, rel=1e-3)   # twice the light time of 1 au
    for p, a, inc, e, w, rs in ((2.0, 8.0, 1.5, 0.0, 0.0, 1.0), (5.0, 14.0, 1.55, 0.3, 0.9, 0.8)):
        te = eo.eclipse_phase(p, inc, e, w)
        los = []
        for t in (0.0, te):
            f = ro.ta_newton(t, 0.0, p, e, w)
            r = a * (1 - e ** 2) / (1 + e * math.cos(f))
            los.append(r * math.sin(w + f) * math.sin(inc))   # distance towards the observer
        expect = (los[0] - los[1]) * rs * eo.R_SUN_M / eo.C_LIGHT / 86400.0
        assert eo.light_travel_time(p, a, inc, e, w, rs) == pytest.approx(expect, rel=1e-7)


def test_eclipse_flux_is_visible_area_and_ltt_shift():
    """In the window the flux is pi k^2 - A(1, k, z) on the eclipse series; rstar
    moves the eclipse by the light travel time."""
    args = list(reference_eclipse_inputs(npt=2001, t_start=0.85, t_end=1.15))
    times = args[0]
    f1 = eo.eclipse_model(*args)[0]
    te, c = eo.taylor_coeffs_eclipse(2.0, 8.0, np.pi / 2, 0.0, 0.0)
    ltt = eo.light_travel_time(2.0, 8.0, np.pi / 2, 0.0, 0.0, 1.0)
    tc = times - (te + ltt)
    k = 0.1
    t1, t4 = eo.bounding_box_eclipse(k, c)
    inw = (t1 - 0.003 <= tc) & (tc <= t4 + 0.003)
    expect = np.where(inw, math.pi * k * k - ro.cci_area(1.0, k, ro.sep(tc, c)), math.pi * k * k)
    np.testing.assert_allclose(f1, expect, atol=1e-13)
    args[8] = 10.0   # a larger star: a longer delay, the same shape shifted
    f10 = eo.eclipse_model(*args)[0]
    shift = eo.light_travel_time(2.0, 8.0, np.pi / 2, 0.0, 0.0, 10.0) - ltt
    mid = lambda f: np.sum(times * (math.pi * k * k - f)) / np.sum(math.pi * k * k - f)
    assert mid(f10) - mid(f1) == pytest.approx(shift, rel=1e-3)


def test_esmodel_is_the_eclipse_model_normalised():
    """esmodel = (1 + f EM / pi) / (1 + f k^2), EM the eclipse model's visible area,
    wherever both windows cover the eclipse; 1 outside."""
    times = np.linspace(0.8, 1.2, 801)
    npv = 3
    k = np.array([0.08, 0.1, 0.12])
    p, a, inc = np.full(npv, 2.0), np.array([7.0, 8.0, 9.0]), np.array([1.53, 1.55, 1.57])
    e, w, rs = np.zeros(npv), np.zeros(npv), np.array([0.9, 1.0, 1.1])
    fr = np.array([[1e-3, 5e-3], [2e-3, 4e-3], [3e-3, 1e-2]])
    es = eo.esmodel(times, k, np.zeros(npv), p, a, inc, e, w, rs, fr, 3, 0.01)
    for ipv in range(npv):
        em = eo.eclipse_model(times, k[ipv:ipv + 1], np.zeros((1, 1)), p[ipv:ipv + 1], a[ipv:ipv + 1],
                              inc[ipv:ipv + 1], e[ipv:ipv + 1], w[ipv:ipv + 1], rs[ipv], 1,
                              np.zeros(times.size, np.int64), np.zeros(1, np.int64), np.array([3]),
                              np.array([0.01]))[0]
        for b in range(2):
            expect = (1 + fr[ipv, b] * em / math.pi) / (1 + fr[ipv, b] * k[ipv] ** 2)
            np.testing.assert_allclose(es[ipv, b], expect, atol=1e-14)
    assert es.min() < 1 - 1e-5


def test_ellipse_rows_lie_on_the_rotated_ellipse():
    """Each row end satisfies (u/k)^2 + (v/(k(1-f)))^2 = 1 with (u, v) the point
    rotated back by -alpha; rows off the ellipse are NaN."""
    for ny, k, f, al in ((100, 0.1, 0.2, 0.5), (57, 0.15, 0.05, -1.2), (200, 0.08, 0.4, 2.0)):
        ys, xs = eo.create_ellipse(ny, k, f, al)
        ok = np.isfinite(xs[:, 0])
        assert ok.sum() > ny // 2
        for j in np.nonzero(ok)[0]:
            for x in xs[j]:
                u = x * math.cos(al) + ys[j] * math.sin(al)
                v = -x * math.sin(al) + ys[j] * math.cos(al)
                assert (u / k) ** 2 + (v / (k * (1 - f))) ** 2 == pytest.approx(1.0, abs=1e-12)
        assert ys[1] - ys[0] == pytest.approx(2 * k / (ny + 1))


def _grid_area(cx, cy, k, f, al, n=4001):
    # brute force: midpoint grid over the ellipse's bounding square
    h = 2 * k / n
    g = -k + h * (np.arange(n) + 0.5)
    X, Y = np.meshgrid(g, g)
    u = X * math.cos(al) + Y * math.sin(al)
    v = -X * math.sin(al) + Y * math.cos(al)
    inside_e = (u / k) ** 2 + (v / (k * (1 - f))) ** 2 <= 1.0
    inside_s = (X + cx) ** 2 + (Y + cy) ** 2 <= 1.0
    return (inside_e & inside_s).sum() * h * h


@pytest.mark.parametrize('cx,cy', [(-0.95, 0.1), (0.97, -0.2), (0.02, 0.99), (-0.5, 0.85), (0.6, -0.75)])
def test_scanline_area_against_brute_force(cx, cy):
    """All three branches (centre left of -k, right of k, between) against a grid."""
    k, f, al = 0.1, 0.25, 0.6
    z = math.hypot(cx, cy)
    ref = _grid_area(cx, cy, k, f, al)      # itself good to ~5e-6 relative
    assert 0.0 < ref < math.pi * k * k * (1 - f)
    # the scanline sum converges to the area as the rows get denser
    for ny, rel in ((400, 3e-4), (4000, 2e-5)):
        ys, xs = eo.create_ellipse(ny, k, f, al)
        assert eo.ellipse_circle_area(cx, cy, z, k, f, xs, ys) == pytest.approx(ref, rel=rel)


def test_scanline_area_limits():
    k, f, al = 0.12, 0.3, 0.4
    ys, xs = eo.create_ellipse(100, k, f, al)
    assert eo.ellipse_circle_area(0.1, 0.2, math.hypot(0.1, 0.2), k, f, xs, ys) == math.pi * k * (1 - f) * k
    assert eo.ellipse_circle_area(1.2, 0.1, math.hypot(1.2, 0.1), k, f, xs, ys) == 0.0
    # f = 0: the circle lens area of rr_oracle, approached as the rows get denser
    # (the error of the row sum falls about as ny^-1.5)
    for cx, cy in ((-0.93, 0.2), (0.5, 0.82), (0.05, -0.97)):
        z = math.hypot(cx, cy)
        lens = float(ro.cci_area(1.0, k, z))
        err = []
        for ny in (4000, 16000):
            ys, xs = eo.create_ellipse(ny, k, 0.0, al)
            err.append(abs(eo.ellipse_circle_area(cx, cy, z, k, 0.0, xs, ys) / lens - 1))
        assert err[0] < 1e-5 and err[1] < err[0] / 4


def test_opmodel_with_zero_flattening_converges_to_roadrunner():
    """f = 0: a circular planet, so the oblate model is rr_full up to the scanline
    discretisation of the limb-crossing area (it shrinks with nlines), with the
    window margin 0.0015 d instead of 0.003 d."""
    om = ro.OracleRR('quadratic')
    times = np.linspace(-0.15, 0.15, 601)
    npt = times.size
    om.set_data(times)
    npv = 3
    k = np.array([[0.08], [0.1], [0.12]])
    ldc = np.array([[[0.4, 0.2]], [[0.3, 0.1]], [[0.5, 0.25]]])
    t0 = np.zeros((npv, 1))
    p, a, inc = np.array([2.0, 3.0, 4.0]), np.array([7.0, 9.0, 11.0]), np.array([1.55, 1.53, 1.56])
    e, w = np.zeros(npv), np.zeros(npv)
    rr = om.evaluate(k, ldc, t0, p, a, inc, e, w)
    ldp, istar = om.ld(ldc)
    errs = []
    for nl in (100, 800):
        op = eo.opmodel(times, k, np.zeros(npv), np.zeros(npv), t0, p, a, inc, e, w, 1, 1, 1, nl,
                        np.zeros(npt, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64), np.ones(1, np.int64),
                        np.zeros(1), ldp, istar, om.weights, om.dk, 0.005, 0.5, om.dg, om.ze)
        errs.append(np.abs(op - rr).max())
    assert errs[1] < 3e-6 and errs[1] < errs[0] / 4
    assert (rr < 1).sum() > 300


def test_opmodel_full_transit_depth_is_ellipse_area():
    """Fully inside the disk the blocked area is pi k^2 (1 - f): with uniform limb
    darkening the depth is k^2 (1 - f) exactly."""
    om = ro.OracleRR('uniform')
    times = np.array([-0.3, -0.01, 0.0, 0.01, 0.3])
    om.set_data(times)
    ldc = np.zeros((1, 1, 0))
    ldp, istar = om.ld(ldc)
    k, f = 0.1, 0.3
    op = eo.opmodel(times, np.array([[k]]), np.array([f]), np.array([0.7]), np.zeros((1, 1)), np.array([3.0]),
                    np.array([9.0]), np.array([0.5 * math.pi]), np.zeros(1), np.zeros(1), 1, 1, 1, 100,
                    np.zeros(5, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64), np.ones(1, np.int64),
                    np.zeros(1), ldp, istar, om.weights, om.dk, 0.005, 0.5, om.dg, om.ze)[0]
    np.testing.assert_allclose(op[1:4], 1.0 - k * k * (1 - f), rtol=0, atol=1e-14)
    assert op[0] == 1.0 and op[4] == 1.0
