"""rr_evaluate_noalloc (workspace in the library's per-GPU scratch array) against
rr_evaluate with a caller-allocated workspace: the same kernels on the same inputs,
so the results must be identical bit for bit.  Two calls in a row, the second with
a smaller problem, so a stale scratch array from the first call would show."""
import ctypes

import numpy as np
import pytest
import torch

from pytransit_amd import RoadRunnerModel, _lib, workloads
from pytransit_amd._lib import check, ptr, stream_handle
from pytransit_amd.rrmodel import _PRECISIONS

pytestmark = pytest.mark.gpu


def _noalloc(tm, args):
    lib = _lib.load()
    npv, tab, lc, prm, ld, keep = tm._descriptors(*args)
    prec = _PRECISIONS[tm.precision]
    dtype = torch.float64 if prec == _lib.RR_FP64 else torch.float32
    flux = torch.full((npv, tm.npt), -7.0, dtype=dtype, device=tm.device)
    byref = ctypes.byref
    check(lib.rr_evaluate_noalloc(prec, byref(tab), byref(lc), byref(prm), byref(ld), ptr(flux),
                                  stream_handle(tm.device)))
    torch.cuda.synchronize()
    return flux


def _same(a, b):
    assert a.shape == b.shape
    assert torch.equal(torch.isnan(a), torch.isnan(b))
    m = ~torch.isnan(a)
    assert torch.equal(a[m], b[m]), float((a[m].double() - b[m].double()).abs().max())


@pytest.mark.parametrize('prec', ['fp64', 'mixed'])
def test_noalloc_equals_rr_evaluate_twice(gpu, prec):
    # call 1: BJD-scale light curve, power-2, nss=10, 512 pv
    wl = workloads.rr_bjd(npv=512, npt=20000)
    tm = RoadRunnerModel(wl.law, precision=prec, device=gpu)
    tm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    args = (wl.k, wl.ldc, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w)
    print(f'noalloc {prec}: call 1, 512 pv x 20000 t', flush=True)
    ref = tm._evaluate_device(*args)
    got = _noalloc(tm, args)
    _same(got, ref)
    assert (ref < 1).any()
    # call 2: a smaller problem with another law and several passbands
    rng = np.random.default_rng(11)
    npv, npt, npb = 40, 3000, 3
    times = np.linspace(-0.4, 0.4, npt)
    lcids = np.repeat(np.arange(npb), npt // npb)
    tm2 = RoadRunnerModel('quadratic', precision=prec, device=gpu)
    tm2.set_data(times, lcids=lcids, pbids=np.arange(npb))
    k = rng.uniform(0.05, 0.15, (npv, npb))
    ldc = np.stack([rng.uniform(0.1, 0.6, (npv, npb)), rng.uniform(0, 0.3, (npv, npb))], -1)
    t0 = rng.normal(0, 0.005, (npv, 1))
    p = rng.uniform(1, 10, npv)
    a = rng.uniform(4, 15, npv)
    inc = np.arccos(rng.uniform(0, 0.9, npv) / a)
    a[3] = 0.5  # an invalid pv gives a NaN row in both
    args2 = (k, ldc, t0, p, a, inc, 0.0, 0.0)
    print(f'noalloc {prec}: call 2, {npv} pv x {npt} t x {npb} pb', flush=True)
    ref2 = tm2._evaluate_device(*args2)
    got2 = _noalloc(tm2, args2)
    _same(got2, ref2)
    assert torch.isnan(ref2[3]).all() and (ref2 < 1).any()


def test_noalloc_rejects_oversize_before_device_work(gpu):
    lib = _lib.load()
    tab = _lib.RRTable(nk=256, ng=100, nz=40)
    lc = _lib.RRLcData(npt=1000, nlc=1, npb=1, nep=1)
    prm = _lib.RRParams(npv=50000, kcols=1, t0cols=1)
    ld = _lib.RRLd(law=2, nldc=2)
    byref = ctypes.byref
    rc = lib.rr_evaluate_noalloc(0, byref(tab), byref(lc), byref(prm), byref(ld), None, None)
    assert rc == -3 and b'scratch' in lib.rr_last_error()
