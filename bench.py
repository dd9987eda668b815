#!/usr/bin/env python3
"""Benchmark of the RoadRunner hot path on one MI355X.

A step is one ``RoadRunnerModel.evaluate`` of the whole parameter population on
device-resident inputs: limb-darkening profiles, per-pv setup (Taylor orbit,
contact window, limb-darkening means) and the flux kernel.  The workload is the
BJD-scale population: 4,096 parameter vectors x 20,000 times (2-min cadence from
BJD 2459000.0), 10 supersamples over a 30-min exposure, power-2 limb darkening.

``value`` is model points/s (one point = one output flux value) of the mixed
precision path (FP64 epoch reduction and setup, FP32 flux).  The same workload is
timed in plain FP64 and the maximum |F_mixed - F_fp64| over every point is
reported in ppm.  ``roofline`` is for the flux kernel alone, timed with HIP events
on the stream it is launched on: algorithmic HBM bytes per launch against 8 TB/s,
and (``roofline.flop``) algorithmic FLOPs per launch, counted exactly by the
oracle on this workload (profiles/rr_flux_rr_bjd_flop_count.md), against the
78.6 TFLOP/s vector peak; ``bound`` names the larger fraction.  ``cpu_baseline``
is the C/OpenMP FP64 restatement of the reference algorithm on this host's cores.

Further legs, reported beside the headline in the same JSON line:
* ``ts_many_pb``: TSModel, 1,024 pv x 5,000 times x 512 passbands (mixed and FP64);
* ``gradient``: fused log-likelihood with its adjoint gradient, 2,048 pv x 10,000
  times (mixed and FP64);
* ``c0``: the CPU-runnable case, 64 pv x 2,000 times, quadratic, nss=1 (FP64,
  mixed, and the C/OpenMP restatement on the host);
* ``with_d2h``: the headline workload with ``copy=True`` (flux copied to host);
* ``serving``: a Poisson stream of 1000 fit requests through the request server,
  batched (4 slots, up to 16 requests and 2^28 points per launch), batched with
  the launches replayed from captured HIP graphs, and unbatched (one at a time),
  with throughput and p50/p99 latency of the second of two passes over the
  stream (the first warms each server); ``capacity`` offers the same mix at
  twice the rate and reports what each serves.

Usage: python bench.py [--gpus 1] [--steps 100] [--warmup 3] [--no-cpu] [--no-serve] [--no-extra]
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

import numpy as np

METRIC = "model points/s mixed vs FP64 plain at ≤1 ppm flux err; served p50/p99 lat"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--cpu-reps', type=int, default=5)
    ap.add_argument('--no-serve', action='store_true', help='skip the serving leg')
    ap.add_argument('--no-extra', action='store_true',
                    help='skip the many-passband TSModel and likelihood-gradient legs')
    ap.add_argument('--serve-requests', type=int, default=1000)
    ap.add_argument('--serve-rate', type=float, default=2000.0, help='Poisson arrival rate, requests/s')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also launch a plain 256 MiB device copy (the calibration kernel of the HBM counter runs)')
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                            'rr_flux_mixed_rr_bjd_hbm_traffic.md')
FLOP_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                         'rr_flux_rr_bjd_flop_count.md')
# MI355X vector peaks without packed-FP32 instructions (the library is built
# without them): 256 CUs x 128 FLOP/clock x 2.4 GHz, for FP32 and FP64 alike.
VALU_PEAK_TFLOPS = 78.6


def measured_traffic():
    """HBM bytes per mixed flux launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r'traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)', f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def flop_count():
    """Algorithmic FLOPs of one flux evaluation of the rr_bjd workload, counted
    with the oracle's window and geometry (profiles/rr_flux_rr_bjd_flop_count.md,
    written by ``python -m oracle.flop_count``), or None."""
    try:
        with open(FLOP_FILE) as f:
            m = re.search(r'flops_per_launch\s*=\s*([0-9]+)', f.read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def copy_probe(torch, dev):
    """A copy of known size: 256 MiB read + 256 MiB written per launch."""
    n = 64 << 20
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def flux_kernel_bytes(npv, npt, ng, out_bytes):
    """Algorithmic HBM bytes of one flux launch: the flux write, one read of the
    times, and per pv its 192-byte setup record (orbit coefficients, window and
    light curve 0's zero epoch, radius ratio and disk integral) and its
    limb-darkening-mean row (ng values)."""
    row = ng * (4 if out_bytes == 4 else 8)
    return npv * npt * out_bytes + npt * 8 + npv * (192 + row)


def time_path(tm, args_dev, steps, warmup):
    """Seconds per evaluation over `steps` back-to-back evaluations, then the mean
    flux-kernel duration (ms) from HIP events around each flux launch of a second
    run of `steps` evaluations.  The two are separate because an event pair around
    every launch adds its own gaps between the kernels of a step."""
    import torch
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for _ in range(warmup):
        tm._evaluate_device(*args_dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for s in range(steps):
        out = tm._evaluate_device(*args_dev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for s in range(steps):
        tm._evaluate_device(*args_dev, flux_events=ev[s])
    torch.cuda.synchronize()
    kern_ms = statistics.mean(a.elapsed_time(b) for a, b in ev)
    return wall / steps, kern_ms, out


def serve_stream(torch, dev, st, slots, max_batch, budget, max_points=None, graphs=False):
    """Replay the request stream in real time through one server, twice: the first
    pass warms the server (every light-curve set resident, kernels loaded and,
    with ``graphs``, the launch graphs of the stream's shapes captured), the second
    is measured.  Returns the second pass's per-request latencies (s), the points
    served, the wall time from the first arrival to the last completion, and the
    graphs captured during it."""
    from pytransit_amd.serving import Dataset, RRServer
    dsets = [Dataset(**d) for d in st.datasets]
    srv = RRServer(hbm_budget=budget, slots=slots, max_batch=max_batch, precision='mixed', device=dev,
                   auto_release=True, max_batch_points=max_points, graphs=graphs)
    reqs = st.requests
    n = len(reqs)
    for npass in range(2):
        torch.cuda.synchronize()
        srv.completed.clear()
        captures0 = srv._exec.graph_captures
        t_start = time.perf_counter()
        i = 0
        while i < n or not srv.idle:
            now = time.perf_counter() - t_start
            while i < n and reqs[i]['arrival'] <= now:
                rq = reqs[i]
                srv.submit(dsets[rq['dataset']], rq['k'], rq['ldc'], rq['t0'], rq['p'], rq['a'], rq['i'],
                           ld=rq['ld'], arrival=t_start + rq['arrival'])
                i += 1
            srv.step()
    done = srv.completed
    lat = np.array([r.latency for r in done])
    wall = max(r.t_done for r in done) - t_start
    points = sum(r.points for r in done)
    return dict(latency=lat, points=points, wall=wall, batch=np.mean([r.batch_size for r in done]),
                captures=srv._exec.graph_captures - captures0, graphs=srv._exec.graph_captures)


def serving_leg(torch, dev, nreq, rate):
    st = workloads_mod().serving_stream(n_requests=nreq, rate=rate)
    budget = 48 << 30
    out = {}
    # batched: up to 16 requests and 2^28 points (~0.6 ms of device time) per launch;
    # batched_graphs: the same with the launches replayed from captured graphs
    servers = (('batched', 4, 16, 1 << 28, False), ('batched_graphs', 4, 16, 1 << 28, True),
               ('unbatched', 1, 1, None, False))
    for name, slots, mb, mp, gr in servers:
        r = serve_stream(torch, dev, st, slots, mb, budget, mp, gr)
        out[name] = {"throughput_points_per_s": r['points'] / r['wall'], "requests_per_s": nreq / r['wall'],
                     "p50_ms": float(np.percentile(r['latency'], 50) * 1e3),
                     "p99_ms": float(np.percentile(r['latency'], 99) * 1e3),
                     "wall_s": r['wall'], "mean_batch": float(r['batch']), "slots": slots, "max_batch": mb,
                     "max_batch_points": mp}
        if gr:
            out[name]["graphs"] = r['graphs']
            out[name]["graphs_captured_in_measured_pass"] = r['captures']
    # capacity: the same mix offered at twice the rate
    st2 = workloads_mod().serving_stream(n_requests=nreq, rate=2 * rate)
    cap = {"offered_requests_per_s": 2 * rate}
    for name, slots, mb, mp, gr in servers:
        r = serve_stream(torch, dev, st2, slots, mb, budget, mp, gr)
        cap[name + "_points_per_s"] = r['points'] / r['wall']
        cap[name + "_requests_per_s"] = nreq / r['wall']
        cap[name + "_p50_ms"] = float(np.percentile(r['latency'], 50) * 1e3)
        cap[name + "_p99_ms"] = float(np.percentile(r['latency'], 99) * 1e3)
    out["capacity"] = cap
    out["workload"] = (f"{nreq} fit requests, Poisson arrivals at {rate:g}/s; light-curve sets of log-uniform "
                       f"1k-200k points in 1-64 passbands (24-set pool, resident), log-uniform 32-8192 pv, "
                       f"quadratic/power-2/nonlinear-4 LD, nss in {{1,5,10}}; mixed precision; flux left on device")
    out["points"] = int(sum(rq['k'].shape[0] * st.datasets[rq['dataset']]['time'].size for rq in st.requests))
    out["hbm_budget_bytes"] = budget
    return out


def _timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / steps, out


def ts_leg(torch, dev, steps, warmup):
    """Many-passband TSModel: 1,024 pv x 5,000 times x 512 passbands, tabulated
    128-node profiles, exact weights at mean(k) per pv."""
    from pytransit_amd import TabulatedLDModel, TSModel, workloads
    probe = TSModel('quadratic', nzin=64, nzlimb=64)
    wl = workloads.ts_many_pb(mu=probe.mu)
    ldm = TabulatedLDModel(torch.as_tensor(wl.ldp, device=dev), torch.as_tensor(wl.istar, device=dev))
    args = tuple(torch.as_tensor(x, device=dev) for x in (wl.k, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w))
    res = {}
    for prec in ('mixed', 'fp64'):
        tm = TSModel(ldm, nzin=64, nzlimb=64, precision=prec, device=dev)
        tm.set_data(wl.times)
        k, t0, p, a, i, e, w = args
        sec, out = _timed(torch, lambda: tm._evaluate_device(k, None, t0, p, a, i, e, w), steps, warmup)
        res[prec] = (sec, out)
    npv, npb, npt = res['mixed'][1].shape
    err = 0.0
    for j in range(0, npv, 64):  # chunked to bound memory
        d = (res['mixed'][1][j:j + 64].double() - res['fp64'][1][j:j + 64]).abs().max()
        err = max(err, float(d))
    pts = npv * npb * npt
    wbytes = pts * 4
    return {"workload": f"TSModel, npv={npv} x npb={npb} x npt={npt}, 128 mu nodes, tabulated LD "
                        f"(LDModel), exact weights at mean(k)",
            "points": pts, "mixed_points_per_s": pts / res['mixed'][0], "fp64_points_per_s": pts / res['fp64'][0],
            "mixed_ms": res['mixed'][0] * 1e3, "fp64_ms": res['fp64'][0] * 1e3,
            "mixed_over_fp64": res['fp64'][0] / res['mixed'][0], "max_abs_err_ppm_mixed_vs_fp64": err * 1e6,
            "flux_write_floor_ms": wbytes / (HBM_PEAK_GBS * 1e9) * 1e3}


def grad_leg(torch, dev, steps, warmup):
    """Log-likelihood and its adjoint gradient: 2,048 pv x 10,000 times, nss=5,
    quadratic LD, 300 ppm noise."""
    from pytransit_amd import RoadRunnerModel, workloads
    wl = workloads.rr_gradient()
    tr = wl.truth
    names = ('k', 't0', 'p', 'a', 'i', 'u1', 'u2')
    theta = torch.as_tensor(np.stack([wl.pop[n] for n in names], 1), device=dev)
    res = {}
    for prec in ('mixed', 'fp64'):
        tm = RoadRunnerModel('quadratic', precision=prec, device=dev)
        tm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
        if prec == 'mixed':
            ref = RoadRunnerModel('quadratic', precision='fp64', device=dev)
            ref.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
            f_true = ref._evaluate_device([[tr['k']]], [[[tr['u1'], tr['u2']]]], [[tr['t0']]], tr['p'], tr['a'],
                                          tr['i'], 0.0, 0.0)[0]
            obs = f_true + torch.as_tensor(wl.noise_draw, device=dev)
        th = theta
        fn = lambda: tm._lnlike_device(th[:, 0], th[:, 5:7].reshape(-1, 1, 2), th[:, 1], th[:, 2], th[:, 3],
                                       th[:, 4], obs, wl.noise, grad=True)
        res[prec] = _timed(torch, fn, steps, warmup)
    npv = theta.shape[0]
    gm, gf = res['mixed'][1][1], res['fp64'][1][1]
    scale = gf.abs().pow(2).mean(0).sqrt()
    rel = ((gm - gf).abs() / (gf.abs() + scale)).max()
    return {"workload": f"lnL + d lnL/d(k,t0,p,a,i,u1,u2), npv={npv} x npt={wl.times.size}, nss=5, quadratic",
            "mixed_ms": res['mixed'][0] * 1e3, "fp64_ms": res['fp64'][0] * 1e3,
            "mixed_points_per_s": npv * wl.times.size / res['mixed'][0],
            "fp64_points_per_s": npv * wl.times.size / res['fp64'][0],
            "max_rel_grad_diff_mixed_vs_fp64": float(rel),
            "max_abs_lnl_diff_mixed_vs_fp64": float((res['mixed'][1][0] - res['fp64'][1][0]).abs().max())}


def c0_leg(torch, dev, steps, warmup, cpu):
    """The CPU-runnable case: 64 pv x 2,000 times, quadratic LD, no supersampling,
    FP64 and mixed, on the device and (``cpu``) with the C/OpenMP restatement.
    It writes 1 MB, so its time is launch latency, not a roofline position."""
    from pytransit_amd import RoadRunnerModel, workloads
    wl = workloads.rr_small()
    npts = wl.k.shape[0] * wl.times.size
    args_dev = tuple(torch.as_tensor(x, device=dev) for x in (wl.k, wl.ldc, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w))
    out = {"workload": f"rr_small: RoadRunnerModel('quadratic') evaluate, npv={wl.k.shape[0]} x npt={wl.times.size}, "
                       f"nss=1", "points": npts}
    for prec in ('fp64', 'mixed'):
        tm = RoadRunnerModel(wl.law, precision=prec, device=dev)
        tm.set_data(wl.times)
        sec, _ = _timed(torch, lambda: tm._evaluate_device(*args_dev), steps, warmup)
        out[prec + "_points_per_s"] = npts / sec
        out[prec + "_ms"] = sec * 1e3
    if cpu:
        from oracle import rr_cpu
        nthreads = min(len(os.sched_getaffinity(0)), int(os.environ.get('OMP_NUM_THREADS', 1 << 30)))
        cm = rr_cpu.CpuRR('quadratic')
        cm.set_data(wl.times)
        a = (wl.k, wl.ldc, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w)
        cm.evaluate(*a, nthreads=nthreads)
        ts = []
        for _ in range(20):
            t = time.perf_counter()
            cm.evaluate(*a, nthreads=nthreads)
            ts.append(time.perf_counter() - t)
        out["cpu_points_per_s"] = npts / statistics.median(ts)
        out["cpu_cores"] = nthreads
    return out


def d2h_rate(torch, tm, args_dev, steps):
    """points/s of evaluate(copy=True): the device result copied to host memory
    (numpy), i.e. the PCIe-inclusive rate of the reference's call semantics."""
    npv, npt = args_dev[0].shape[0], tm.npt
    tm.evaluate(*args_dev, copy=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        f = tm.evaluate(*args_dev, copy=True)
    sec = (time.perf_counter() - t) / steps
    return {"points_per_s": npv * npt / sec, "ms": sec * 1e3, "bytes": int(f.nbytes)}


def workloads_mod():
    from pytransit_amd import workloads
    return workloads


def cpu_baseline(wl, reps):
    from oracle import rr_cpu
    ncores = len(os.sched_getaffinity(0))
    nthreads = min(ncores, int(os.environ.get('OMP_NUM_THREADS', ncores)))
    cm = rr_cpu.CpuRR('power-2')
    cm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
    args = (wl.k, wl.ldc, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w)
    cm.evaluate(*args, nthreads=nthreads)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        cm.evaluate(*args, nthreads=nthreads)
        ts.append(time.perf_counter() - t)
    med = statistics.median(ts)
    npts = wl.k.shape[0] * wl.times.size
    # single-thread figure on a 256-pv slice of the same workload
    n1 = 256
    a1 = tuple(x[:n1] for x in args)
    t = time.perf_counter()
    cm.evaluate(*a1, nthreads=1)
    t1 = time.perf_counter() - t
    cpu_model = ''
    try:
        with open('/proc/cpuinfo') as f:
            cpu_model = next(l.split(':', 1)[1].strip() for l in f if l.startswith('model name'))
    except (OSError, StopIteration):
        pass
    return {"value": npts / med, "unit": "points/s", "cores": nthreads, "kind": "port",
            "sample": f"full workload ({wl.k.shape[0]}x{wl.times.size} points, nss=10, power-2), C/OpenMP FP64 "
                      f"restatement of rr_full, median of {reps} runs after 1 warm-up, LD evaluation included",
            "value_1thread": n1 * wl.times.size / t1, "sample_1thread": f"{n1}x{wl.times.size} points, 1 thread",
            "cpu_model": cpu_model}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path has no multi-GPU form "
              f"in this build.", file=sys.stderr)
        sys.exit(2)
    import torch
    from pytransit_amd import RoadRunnerModel, workloads

    dev = torch.device('cuda', 0)
    if args.copy_probe:
        copy_probe(torch, dev)
    wl = workloads.rr_bjd()
    npv, npt = wl.k.shape[0], wl.times.size
    args_dev = tuple(torch.as_tensor(x, device=dev) for x in (wl.k, wl.ldc, wl.t0, wl.p, wl.a, wl.i, wl.e, wl.w))

    res = {}
    d2h = {}
    for prec in ('mixed', 'fp64'):
        tm = RoadRunnerModel(wl.law, precision=prec, device=dev)
        tm.set_data(wl.times, nsamples=wl.nsamples, exptimes=wl.exptimes)
        step_s, kern_ms, out = time_path(tm, args_dev, args.steps, args.warmup)
        res[prec] = dict(step_s=step_s, kern_ms=kern_ms, out=out, ng=tm.ng)
        d2h[prec] = d2h_rate(torch, tm, args_dev, max(3, args.steps // 10))

    err = (res['mixed']['out'].double() - res['fp64']['out']).abs()
    max_err_ppm = float(err.max()) * 1e6
    both_nan = bool(torch.equal(torch.isnan(res['mixed']['out']), torch.isnan(res['fp64']['out'])))

    npts = npv * npt
    mixed_pps = npts / res['mixed']['step_s']
    fp64_pps = npts / res['fp64']['step_s']
    kbytes = flux_kernel_bytes(npv, npt, res['mixed']['ng'], 4)
    achieved = kbytes / (res['mixed']['kern_ms'] * 1e-3) / 1e9
    traffic = measured_traffic()
    flops = flop_count()
    flop = None
    if flops is not None:
        tf = {prec: flops / (res[prec]['kern_ms'] * 1e-3) / 1e12 for prec in res}
        flop = {"achieved": tf['mixed'], "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf['mixed'] / VALU_PEAK_TFLOPS, "flops_per_launch": flops,
                "fp64_achieved": tf['fp64'], "fp64_frac": tf['fp64'] / VALU_PEAK_TFLOPS,
                "peak_note": "vector FP32 without packed instructions, and FP64: 78.6 TFLOP/s",
                "count_source": os.path.relpath(FLOP_FILE)}
    hbm_frac = achieved / HBM_PEAK_GBS
    flop_frac = flop['frac'] if flop else 0.0
    line = {
        "metric": METRIC,
        "value": mixed_pps,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res['mixed']['step_s'] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"rr_bjd: RoadRunnerModel('power-2') evaluate, npv={npv} x npt={npt} "
                               f"(2-min cadence from BJD 2459000.0), nss=10 over 30 min, mixed precision",
                   "npv": npv, "npt": npt, "nsamples": 10, "ld": "power-2"},
        "fp64_points_per_s": fp64_pps,
        "fp64_ms_per_step": res['fp64']['step_s'] * 1e3,
        "mixed_over_fp64": mixed_pps / fp64_pps,
        "max_abs_err_ppm_mixed_vs_fp64": max_err_ppm,
        "nan_pattern_equal": both_nan,
        "roofline": {"bound": "hbm" if hbm_frac >= flop_frac else "valu", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_frac, "traffic": traffic,
                     "kernel": "rr_flux_kernel<mixed>", "kernel_ms": res['mixed']['kern_ms'],
                     "algorithmic_bytes": kbytes,
                     "fp64_kernel_ms": res['fp64']['kern_ms'],
                     "traffic_source": os.path.relpath(TRAFFIC_FILE) if traffic is not None else None,
                     "flop": flop},
        "with_d2h": {"mixed_points_per_s": d2h['mixed']['points_per_s'], "mixed_ms": d2h['mixed']['ms'],
                     "fp64_points_per_s": d2h['fp64']['points_per_s'], "fp64_ms": d2h['fp64']['ms'],
                     "note": "evaluate(copy=True): the flux copied to a numpy array each step"},
    }
    line["c0"] = c0_leg(torch, dev, args.steps, args.warmup, not args.no_cpu)
    if not args.no_extra:
        line["ts_many_pb"] = ts_leg(torch, dev, max(3, args.steps // 4), 1)
        line["gradient"] = grad_leg(torch, dev, max(3, args.steps // 4), 1)
    if not args.no_serve:
        srv = serving_leg(torch, dev, args.serve_requests, args.serve_rate)
        line["serving"] = srv
        line["served_p50_ms"] = srv["batched"]["p50_ms"]
        line["served_p99_ms"] = srv["batched"]["p99_ms"]
    if not args.no_cpu:
        cb = cpu_baseline(wl, args.cpu_reps)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = mixed_pps / cb["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
