"""Algorithmic FLOPs of one RoadRunner flux evaluation, counted exactly with the
FP64 oracle's window and geometry.

TEST INFRASTRUCTURE (part of the oracle): run by hand, not by the product path.
It writes ``profiles/rr_flux_rr_bjd_flop_count.md``, whose totals ``bench.py``
reads to put the flux kernel's FLOP rate beside its HBM rate.

Convention: add, sub, mul and compare count 1, FMA 2, and div, sqrt, floor and
atan2 1 each.  Per model point (pv, t):

    F = F_phase + sum over its in-window samples of F_sample

* F_phase = 9 (FP64): t - t0, + p/2, / p, floor, the FMA of t0 + epoch p, the
  subtraction for tc, and the two window compares (model_full.py:88-91).
* F_sample = 22 (z: offset add, two 4-term Horner FMA chains, x^2 + y^2, sqrt)
  + 1 (g = z / (1 + k)) + 6 (the g lerp, common.py:225-233) + 2 (accumulate)
  + an area term by branch of the lens area (common.py:52-73):
  2 when the disks do not overlap (z >= 1 + k: one add, one compare),
  3 when the planet is fully inside the disk (z <= 1 - k: compares and pi k^2),
  40 on the limb (the sort of the three lengths, the kite square root, the two
  atan2 angles and the lens sum).

The in-window indicator, the number of samples and each sample's branch come
from the oracle's own phase reduction, contact window (taylor_coeffs,
bounding_box) and Taylor z (sep) on the workload's inputs, so the counts are
exact for the workload, not estimated.  The mixed path computes F_phase in FP64
and F_sample in FP32; the FP64 path computes both in FP64.

Usage: python -m oracle.flop_count [--npv N]   (default: the full 4096 x 20000 workload)
"""
import argparse
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle.rr_oracle import bounding_box, sep, taylor_coeffs  # noqa: E402

F_PHASE = 9
F_Z, F_G, F_LERP, F_ACC = 22, 1, 6, 2
F_AREA = {'outside': 2, 'inside': 3, 'limb': 40}
OUT = os.path.join(ROOT, 'profiles', 'rr_flux_rr_bjd_flop_count.md')


def count(wl, npv=None):
    """Points, in-window points and samples by area branch for workload ``wl``
    (one light curve, uniform nsamples and exposure)."""
    times = np.asarray(wl.times, float)
    nss = int(wl.nsamples)
    exptime = float(wl.exptimes)
    n = wl.k.shape[0] if npv is None else npv
    offs = exptime * ((np.arange(1, nss + 1) - 0.5) / nss - 0.5)
    tot = dict(points=n * times.size, window_points=0, outside=0, inside=0, limb=0)
    for ipv in range(n):
        a, e = float(wl.a[ipv]), float(wl.e[ipv])
        if math.isnan(a) or a <= 1.0 or e < 0.0:
            continue
        k = float(wl.k[ipv, 0])
        p = float(wl.p[ipv])
        c = taylor_coeffs(p, a, float(wl.i[ipv]), e, float(wl.w[ipv]))
        bt1, bt4 = bounding_box(k, c)
        lo, hi = bt1 - (0.003 + exptime), bt4 + (0.003 + exptime)
        t0 = float(wl.t0[ipv, 0])
        epoch = np.floor((times - t0 + 0.5 * p) / p)
        tc = times - (t0 + epoch * p)
        sel = tc[(lo <= tc) & (tc <= hi)]
        tot['window_points'] += sel.size
        if sel.size == 0:
            continue
        z = sep(sel[:, None] + offs[None, :], c)
        out = 1.0 + k <= z
        ins = ~out & (z <= 1.0 - k)
        tot['outside'] += int(out.sum())
        tot['inside'] += int(ins.sum())
        tot['limb'] += int((~out & ~ins).sum())
    return tot


def flops(tot):
    samples = tot['outside'] + tot['inside'] + tot['limb']
    f_phase = F_PHASE * tot['points']
    f_sample = (F_Z + F_G + F_LERP + F_ACC) * samples + sum(F_AREA[b] * tot[b] for b in F_AREA)
    return f_phase, f_sample


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--npv', type=int, default=None, help='count the first N pv only')
    ap.add_argument('--no-write', action='store_true')
    args = ap.parse_args()
    from pytransit_amd import workloads
    wl = workloads.rr_bjd()
    t = time.perf_counter()
    tot = count(wl, args.npv)
    f_phase, f_sample = flops(tot)
    samples = tot['outside'] + tot['inside'] + tot['limb']
    npv = wl.k.shape[0] if args.npv is None else args.npv
    lines = [
        '# Algorithmic FLOPs of the rr_bjd flux evaluation',
        '',
        f'Workload: rr_bjd, {npv} pv x {wl.times.size} times (2-min cadence from BJD 2459000.0), '
        f'nss = {wl.nsamples} over 30 min, power-2 limb darkening (pytransit_amd.workloads.rr_bjd, seed 1).',
        '',
        'Command: `python -m oracle.flop_count` (counts with the FP64 oracle\'s phase reduction, contact window '
        'and Taylor z; convention in the script\'s docstring).',
        '',
        '| quantity | value |',
        '|---|---|',
        f'| model points | {tot["points"]} |',
        f'| points in the window | {tot["window_points"]} ({tot["window_points"] / tot["points"]:.4f}) |',
        f'| in-window samples | {samples} |',
        f'| samples outside the disk | {tot["outside"]} ({tot["outside"] / max(samples, 1):.4f}) |',
        f'| samples fully inside the disk | {tot["inside"]} ({tot["inside"] / max(samples, 1):.4f}) |',
        f'| samples on the limb | {tot["limb"]} ({tot["limb"] / max(samples, 1):.4f}) |',
        f'| F_phase total (FP64) | {f_phase} |',
        f'| F_sample total (FP32 in the mixed path) | {f_sample} |',
        '',
        f'flops_phase_per_launch = {f_phase}',
        f'flops_sample_per_launch = {f_sample}',
        f'flops_per_launch = {f_phase + f_sample}',
        '',
    ]
    text = '\n'.join(lines)
    print(text)
    print(f'[{time.perf_counter() - t:.1f} s]', file=sys.stderr)
    if not args.no_write and args.npv is None:
        with open(OUT, 'w') as f:
            f.write(text)


if __name__ == '__main__':
    main()
