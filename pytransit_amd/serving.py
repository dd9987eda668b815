"""A request server for RoadRunner model evaluations on one GPU.

The reference evaluates one model at a time on the caller's thread
(``RoadRunnerModel.evaluate``, rrmodel.py:175-238); there is no serving layer.
This module adds one for many independent fits sharing a GPU:

* **Resident data.**  The weight table is built once per server.  Light-curve
  sets (``Dataset``: what ``set_data`` takes, transitmodel.py:56-125) are uploaded
  once and stay in device memory while requests use them; idle ones are evicted,
  least recently used first, when memory is needed.
* **HBM-budget admission.**  All device memory of requests and datasets comes
  from one arena of ``hbm_budget`` bytes, sub-allocated first-fit.  A request is
  admitted when a launch slot is free and its parameters, workspace and flux
  output fit; requests that do not fit wait while later, smaller ones are
  admitted, unless the oldest has been passed over ``max_skips`` times, in which
  case admission stops until it fits.
* **Continuous batching.**  Each scheduler step reaps finished batches, frees
  their memory, then admits up to ``max_batch`` waiting requests into one batch on
  a free slot (a HIP stream): one ``rr_evaluate_batch`` call, i.e. one setup launch
  and at most two flux launches for all of them.  Requests that arrive while
  other batches run go into the next batch on another slot, so their work
  overlaps on the device.
* **Replayed launches** (``graphs=True``).  Each slot keeps its batch
  descriptor in a device buffer; a batch is planned on the host
  (``rr_batch_plan``), its descriptor goes to the device with one asynchronous
  copy, and the setup and flux launches (``rr_batch_launch``, which read the
  descriptor from device memory) are a HIP graph captured once per slot and
  launch shape and replayed for every later batch of that shape.  Shapes are
  bucketed (block counts to the next power of two of at least 1024, dynamic LDS
  to 8 KB), blocks past a batch's own exit at once, so few graphs serve a whole
  stream.
* **Latency.**  Each request records its arrival, admission and completion times
  (host clock; completion is when the scheduler sees the batch's event).

A batch evaluates each request with the same kernels and arithmetic as
``RoadRunnerModel.evaluate`` on its own, so results do not depend on the mix.
"""
import time
from collections import OrderedDict, deque
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib
from ._lib import LD_LAW_IDS, RRBatchShape, RRLcData, RRLd, RRParams, RRRequest, check
from .rrmodel import _PRECISIONS, RoadRunnerModel
from .transitmodel import TransitModel, layout_flags

__all__ = ['FreeList', 'Dataset', 'Request', 'RRServer']

ALIGN = 256


def _align(n: int) -> int:
    return (n + ALIGN - 1) // ALIGN * ALIGN


class FreeList:
    """First-fit allocator over byte offsets [0, capacity), ALIGN-aligned blocks,
    with coalescing of neighbouring free blocks."""

    def __init__(self, capacity: int):
        self.capacity = capacity // ALIGN * ALIGN
        self._free = [(0, self.capacity)]      # sorted (offset, size)
        self._used = {}                        # offset -> size
        self.in_use = 0

    def alloc(self, nbytes: int) -> Optional[int]:
        n = _align(max(int(nbytes), 1))
        for j, (off, size) in enumerate(self._free):
            if size >= n:
                if size == n:
                    del self._free[j]
                else:
                    self._free[j] = (off + n, size - n)
                self._used[off] = n
                self.in_use += n
                return off
        return None

    def free(self, off: int) -> None:
        n = self._used.pop(off)
        self.in_use -= n
        lo, hi = 0, len(self._free)
        while lo < hi:
            mid = (lo + hi) // 2
            if self._free[mid][0] < off:
                lo = mid + 1
            else:
                hi = mid
        self._free.insert(lo, (off, n))
        # merge with the right then the left neighbour
        if lo + 1 < len(self._free) and off + n == self._free[lo + 1][0]:
            self._free[lo] = (off, n + self._free[lo + 1][1])
            del self._free[lo + 1]
        if lo > 0 and self._free[lo - 1][0] + self._free[lo - 1][1] == off:
            self._free[lo - 1] = (self._free[lo - 1][0], self._free[lo - 1][1] + self._free[lo][1])
            del self._free[lo]

    def largest_free(self) -> int:
        return max((s for _, s in self._free), default=0)


class _Layout:
    """Offsets of named arrays packed into one ALIGN-aligned block."""

    def __init__(self, parts):
        self.off, self.nbytes = {}, 0
        for name, nbytes in parts:
            self.off[name] = self.nbytes
            self.nbytes += _align(nbytes)


class Dataset:
    """Light curves shared by requests: the arguments of ``set_data``, validated
    as TransitModel.set_data does (transitmodel.py:56-125), kept on the host and
    uploaded to the server's arena on first use."""

    def __init__(self, time, lcids=None, pbids=None, nsamples=None, exptimes=None, epids=None):
        tm = TransitModel()
        tm.set_data(time, lcids, pbids, nsamples, exptimes, epids)
        nsam = tm._per_lc(tm.nsamples, np.int64)
        if (nsam < 1).any():
            raise ValueError('nsamples must be >= 1')
        self.npt, self.nlc, self.npb = tm.npt, tm.nlc, tm.npb
        self.nep = int(np.unique(tm.epids).size) if tm.epids.size else 1
        self.max_epid = int(tm.epids.max()) if tm.epids.size else 0
        self.sorted = layout_flags(tm.time, tm.lcids, tm.epids, tm._per_lc(tm.exptimes, np.float64))
        self.uniform_ns = int(nsam[0]) if (nsam == nsam[0]).all() else 0
        self.host = OrderedDict(times=np.ascontiguousarray(tm.time, np.float64),
                                lcids=tm.lcids.astype(np.int32), pbids=tm.pbids.astype(np.int32),
                                epids=tm.epids.astype(np.int32), nsamples=nsam.astype(np.int32),
                                exptimes=tm._per_lc(tm.exptimes, np.float64))
        self.layout = _Layout([(k, v.nbytes) for k, v in self.host.items()])
        self.nbytes = self.layout.nbytes
        # residency (managed by the server)
        self.offset: Optional[int] = None
        self.ready: Optional[torch.cuda.Event] = None
        self.users = 0
        self.last_used = 0.0


class Request:
    """One model evaluation: a parameter population on a dataset.  ``flux`` is the
    [npv, npt] device result once ``done``; ``release()`` returns its memory."""

    def __init__(self, rid, dataset: Dataset, ld: str, k, ldc, t0, p, a, i, e, w, arrival: float):
        self.rid, self.dataset, self.ld, self.arrival = rid, dataset, ld, arrival
        law, ncoef = LD_LAW_IDS[ld]
        p = np.atleast_1d(np.asarray(p, np.float64))
        npv = p.size
        k = np.asarray(k, np.float64)
        k = k.reshape(npv, -1) if k.size != 1 else np.full((npv, 1), float(k.reshape(-1)[0]))
        if k.shape[1] not in (1, dataset.npb):
            raise ValueError('Radius ratios should be given either as an [npv, 1] or [npv, npb] array.')
        t0 = np.asarray(t0, np.float64)
        t0 = t0.reshape(npv, -1) if t0.size != 1 else np.full((npv, 1), float(t0.reshape(-1)[0]))
        if t0.shape[1] <= dataset.max_epid:
            raise ValueError(f't0 has {t0.shape[1]} columns but the epoch indices go up to {dataset.max_epid}')
        c = np.asarray(ldc, np.float64)
        c = c.reshape((1,) * (3 - c.ndim) + c.shape)
        if c.shape[0] not in (1, npv) or c.shape[1] not in (1, dataset.npb) or c.shape[2] < ncoef:
            raise ValueError(f"ldc of shape {c.shape} does not fit npv={npv}, npb={dataset.npb} and the "
                             f"'{ld}' law ({ncoef} coefficients)")
        c = np.ascontiguousarray(np.broadcast_to(c, (npv, dataset.npb, c.shape[2])))
        vec = [np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64).reshape(-1), (npv,)))
               for x in (p, a, i, e, w)]
        self.npv, self.law, self.nldc = npv, law, c.shape[2]
        self.kcols, self.t0cols = k.shape[1], t0.shape[1]
        self.host = OrderedDict(k=np.ascontiguousarray(k), t0=np.ascontiguousarray(t0), p=vec[0], a=vec[1],
                                inc=vec[2], e=vec[3], w=vec[4], ldc=c)
        self.params = _Layout([(n, v.nbytes) for n, v in self.host.items()])
        self.ws_bytes = 0
        self.flux_bytes = 0
        self.offset: Optional[int] = None        # params + workspace block
        self.flux_offset: Optional[int] = None
        self.flux: Optional[torch.Tensor] = None
        self.t_admit = self.t_done = None
        self.batch_size = 0
        self.skips = 0
        self.done = False
        self._server = None

    @property
    def points(self) -> int:
        return self.npv * self.dataset.npt

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.arrival

    def result(self, copy: bool = True):
        if not self.done:
            raise RuntimeError('request not finished')
        return _lib.to_host(self.flux) if copy else self.flux

    def release(self) -> None:
        if self._server is not None and self.flux_offset is not None:
            self._server._arena.free(self.flux_offset)
            self.flux_offset = None
            self.flux = None


class _Batch:
    def __init__(self, slot, reqs, event):
        self.slot, self.reqs, self.event = slot, reqs, event


class RRServer:
    """Continuous-batching server of RoadRunner evaluations on one device.

    Parameters
    ----------
    hbm_budget : bytes of device memory for datasets, parameters, workspaces and
        flux outputs (one arena, allocated at construction).
    slots : batches in flight at once (one HIP stream each, default priority).
    max_batch : requests per batch (at most ``rr_max_batch()`` of the library).
    max_batch_points : model points per batch (None: no limit).  A request above
        it runs alone; smaller ones are batched up to it, so they do not wait for
        a large request that happens to share their launch.
    precision : 'mixed' or 'fp64', as RoadRunnerModel.
    on_done : called with each finished request; with ``auto_release`` its flux
        memory is returned right after the call.
    graphs : replay each slot's setup and flux launches from HIP graphs captured
        once per launch shape, with the batch descriptor in device memory.
    Table geometry keywords (klims, nk, nzin, nzlimb, zcut, ng) are
    RoadRunnerModel's.
    """

    def __init__(self, hbm_budget: int = 16 << 30, slots: int = 4, max_batch: int = 16, precision: str = 'mixed',
                 device=None, on_done: Optional[Callable[[Request], None]] = None, auto_release: bool = False,
                 max_skips: int = 8, staging_bytes: int = 64 << 20, klims=(0.005, 0.5), nk=256, nzin=20,
                 nzlimb=20, zcut=0.7, ng=100, executor=None, max_batch_points: Optional[int] = None,
                 graphs: bool = False):
        self.precision = precision
        self.max_batch_points = max_batch_points
        self._prec = _PRECISIONS[precision]
        self.max_skips = max_skips
        self.on_done, self.auto_release = on_done, auto_release
        self._arena = FreeList(hbm_budget)
        self._queue: deque = deque()
        self._batches: List[_Batch] = []
        self._datasets: List[Dataset] = []
        self._rid = 0
        self.completed: List[Request] = []
        self._exec = executor if executor is not None else _HipExecutor(self, hbm_budget, slots, max_batch,
                                                                         precision, device, staging_bytes,
                                                                         dict(klims=klims, nk=nk, nzin=nzin,
                                                                              nzlimb=nzlimb, zcut=zcut, ng=ng),
                                                                         graphs)
        self.slots = slots
        self.max_batch = min(max_batch, self._exec.max_batch)
        self._free_slots = deque(range(slots))

    # -- client side ----------------------------------------------------------------
    def submit(self, dataset: Dataset, k, ldc, t0, p, a, i, e=0.0, w=0.0, ld: str = 'quadratic',
               arrival: Optional[float] = None) -> Request:
        """Queue one evaluation of ``evaluate(k, ldc, t0, p, a, i, e, w)`` on
        ``dataset`` with the named limb-darkening law; returns its Request."""
        if ld not in LD_LAW_IDS:
            raise KeyError(ld)
        r = Request(self._rid, dataset, ld, k, ldc, t0, p, a, i, e, w,
                    time.perf_counter() if arrival is None else arrival)
        self._rid += 1
        r._server = self
        r.ws_bytes = self._exec.workspace_bytes(r)
        r.flux_bytes = r.points * (4 if self.precision == 'mixed' else 8)
        need = _align(r.params.nbytes + r.ws_bytes) + _align(r.flux_bytes) + _align(dataset.nbytes)
        if need > self._arena.capacity:
            raise ValueError(f'request {r.rid} needs {need} bytes of device memory; the budget is '
                             f'{self._arena.capacity}')
        if self._exec.staging_need(r) > self._exec.staging_bytes:
            raise ValueError(f'request {r.rid} has {self._exec.staging_need(r)} bytes of parameters; the staging '
                             f'buffer holds {self._exec.staging_bytes}')
        if dataset not in self._datasets:
            self._datasets.append(dataset)
        self._queue.append(r)
        return r

    @property
    def idle(self) -> bool:
        return not self._queue and not self._batches

    def run(self, timeout: Optional[float] = None) -> None:
        """Step until every submitted request has finished."""
        t_end = None if timeout is None else time.perf_counter() + timeout
        while not self.idle:
            if self.step() == 0 and not self._batches and self._queue:
                raise RuntimeError(
                    f'request {self._queue[0].rid} cannot be admitted: nothing is in flight and the arena '
                    f'({self._arena.in_use} of {self._arena.capacity} bytes in use) is held by results that '
                    f'have not been released')
            if t_end is not None and time.perf_counter() > t_end:
                raise TimeoutError('server did not drain in time')

    # -- scheduler ------------------------------------------------------------------
    def step(self) -> int:
        """Reap finished batches, then admit and launch one batch per free slot.
        Returns the number of requests launched."""
        self._reap()
        launched = 0
        while self._free_slots and self._queue:
            reqs = self._admit()
            if not reqs:
                break
            slot = self._free_slots.popleft()
            now = time.perf_counter()
            for r in reqs:
                r.t_admit = now
                r.batch_size = len(reqs)
            event = self._exec.launch(slot, reqs)
            self._batches.append(_Batch(slot, reqs, event))
            launched += len(reqs)
        return launched

    def _reap(self) -> None:
        keep = []
        for b in self._batches:
            if not self._exec.finished(b.event):
                keep.append(b)
                continue
            now = time.perf_counter()
            for r in b.reqs:
                r.t_done = now
                r.done = True
                self._arena.free(r.offset)
                r.offset = None
                r.dataset.users -= 1
                r.dataset.last_used = now
                self.completed.append(r)
                if self.on_done is not None:
                    self.on_done(r)
                if self.auto_release:
                    r.release()
            self._free_slots.append(b.slot)
        self._batches = keep

    def _alloc(self, nbytes: int, keep: Dataset) -> Optional[int]:
        """First-fit block of ``nbytes``; while none fits, evict the least recently
        used idle dataset other than ``keep``.  Returns None when nothing is left to
        evict.  A victim's upload may still be in flight on the upload stream, and
        its block can be refilled at once by a slot stream, so the victim's upload
        is waited for before its block is freed."""
        off = self._arena.alloc(nbytes)
        while off is None:
            idle = [d for d in self._datasets if d.offset is not None and d.users == 0 and d is not keep]
            if not idle:
                return None
            victim = min(idle, key=lambda d: d.last_used)
            self._exec.retire_dataset(victim)
            self._arena.free(victim.offset)
            victim.offset = None
            victim.ready = None
            off = self._arena.alloc(nbytes)
        return off

    def _ensure_resident(self, ds: Dataset) -> bool:
        if ds.offset is not None:
            return True
        off = self._alloc(ds.nbytes, ds)
        if off is None:
            return False
        ds.offset = off
        self._exec.upload_dataset(ds)
        return True

    def _place(self, r: Request) -> bool:
        """Reserve the request's memory (dataset, params + workspace, flux),
        evicting idle datasets other than the request's own when needed."""
        if not self._ensure_resident(r.dataset):
            return False
        off = self._alloc(r.params.nbytes + r.ws_bytes, r.dataset)
        if off is None:
            return False
        foff = self._alloc(r.flux_bytes, r.dataset)
        if foff is None:
            self._arena.free(off)
            return False
        r.offset, r.flux_offset = off, foff
        r.dataset.users += 1
        return True

    def _admit(self) -> List[Request]:
        """Up to max_batch waiting requests in arrival order.  A request that does
        not fit is passed over; once one has been passed over max_skips times, no
        later request is admitted until it fits.  A request that would take the
        batch above max_batch_points waits, without counting as passed over, for
        a later batch, which it heads."""
        batch, rest, passed = [], deque(), []
        staged = points = 0
        blocked = False
        while self._queue:
            r = self._queue.popleft()
            if blocked or len(batch) >= self.max_batch:
                rest.append(r)
                continue
            if batch and self.max_batch_points is not None and points + r.points > self.max_batch_points:
                rest.append(r)
                continue
            size = self._exec.staging_need(r)
            if staged + size <= self._exec.staging_bytes and self._place(r):
                batch.append(r)
                staged += size
                points += r.points
                for q in passed:
                    q.skips += 1
                passed = []
                continue
            rest.append(r)
            passed.append(r)
            if r.skips >= self.max_skips:
                blocked = True
        self._queue = rest
        return batch


class _HipExecutor:
    """Device side of the server: arena tensor, weight table, slot streams, pinned
    staging and the rr_evaluate_batch launches."""

    def __init__(self, server, hbm_budget, slots, max_batch, precision, device, staging_bytes, table_kw,
                 graphs=False):
        self.lib = _lib.load()
        self.server = server
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        self.max_batch = min(max_batch, self.lib.rr_max_batch())
        self.model = RoadRunnerModel('quadratic', precision=precision, device=self.device, **table_kw)
        self.table, self._table_keep = self.model._table(self.device)
        self.arena = torch.empty(server._arena.capacity, dtype=torch.uint8, device=self.device)
        self.base = self.arena.data_ptr()
        self.streams = [torch.cuda.Stream(self.device) for _ in range(slots)]
        self.upload_stream = torch.cuda.Stream(self.device)
        self.staging_bytes = staging_bytes
        self.staging = [torch.empty(staging_bytes, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        self.upload_staging = None
        self.ng = self.model.ng
        self.graphs = graphs
        self.graph_captures = 0
        self.graph_replays = 0
        if graphs:
            nb = int(self.lib.rr_batch_desc_bytes())
            self.desc_host = [torch.empty(nb, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
            self.desc_dev = [torch.empty(nb, dtype=torch.uint8, device=self.device) for _ in range(slots)]
            self.slot_graphs = [{} for _ in range(slots)]   # bucketed shape -> graph, kept for the server's life
            self.capture_stream = torch.cuda.Stream(self.device)

    def workspace_bytes(self, r: Request) -> int:
        d = r.dataset
        return int(self.lib.rr_workspace_bytes(r.npv, d.npt, d.npb, self.ng, d.nlc))

    def staging_need(self, r: Request) -> int:
        return r.params.nbytes

    def _h2d(self, stream, staging: torch.Tensor, soff: int, dst_off: int, arr: np.ndarray) -> int:
        n = arr.nbytes
        if n == 0:
            return soff
        src = staging[soff:soff + n]
        src.numpy()[:] = arr.reshape(-1).view(np.uint8)
        with torch.cuda.stream(stream):
            self.arena[dst_off:dst_off + n].copy_(src, non_blocking=True)
        return soff + _align(n)

    def upload_dataset(self, ds: Dataset) -> None:
        # a fresh pinned buffer per upload: it must outlive the asynchronous copy,
        # which the ready event tracks
        buf = torch.empty(ds.nbytes, dtype=torch.uint8, pin_memory=True)
        soff = 0
        for name, arr in ds.host.items():
            self._h2d(self.upload_stream, buf, ds.layout.off[name], ds.offset + ds.layout.off[name], arr)
        ds.ready = torch.cuda.Event()
        ds.ready.record(self.upload_stream)
        ds._staging = buf

    @staticmethod
    def retire_dataset(ds: Dataset) -> None:
        """Before an evicted dataset's block is reused: its upload must be complete."""
        if ds.ready is not None:
            ds.ready.synchronize()
        ds._staging = None

    def launch(self, slot: int, reqs: List[Request]):
        stream, staging = self.streams[slot], self.staging[slot]
        lib = self.lib
        soff = 0
        descs = (RRRequest * len(reqs))()
        waited = set()
        for j, r in enumerate(reqs):
            ds = r.dataset
            if id(ds) not in waited:
                stream.wait_event(ds.ready)
                waited.add(id(ds))
            base = r.offset
            for name, arr in r.host.items():
                soff = self._h2d(stream, staging, soff, base + r.params.off[name], arr)
            ptrs = {n: self.base + base + o for n, o in r.params.off.items()}
            dptr = {n: self.base + ds.offset + o for n, o in ds.layout.off.items()}
            r.flux = self.arena[r.flux_offset:r.flux_offset + r.flux_bytes].view(
                torch.float32 if self.server.precision == 'mixed' else torch.float64).view(r.npv, ds.npt)
            d = descs[j]
            d.lc = RRLcData(npt=ds.npt, nlc=ds.nlc, npb=ds.npb, nep=r.t0cols, sorted=int(ds.sorted),
                            uniform_nsamples=ds.uniform_ns, times=dptr['times'], lcids=dptr['lcids'],
                            pbids=dptr['pbids'], epids=dptr['epids'], nsamples=dptr['nsamples'],
                            exptimes=dptr['exptimes'])
            d.pv = RRParams(npv=r.npv, kcols=r.kcols, t0cols=r.t0cols, pad_=0, k=ptrs['k'], t0=ptrs['t0'],
                            p=ptrs['p'], a=ptrs['a'], inc=ptrs['inc'], e=ptrs['e'], w=ptrs['w'])
            d.ld = RRLd(law=r.law, nldc=r.nldc, ldc=ptrs['ldc'], ldp=0, istar=0)
            d.workspace = self.base + base + r.params.nbytes
            d.workspace_bytes = r.ws_bytes
            d.flux = self.base + r.flux_offset
        if self.graphs:
            self._replay(slot, stream, reqs, descs)
        else:
            check(lib.rr_evaluate_batch(self.server._prec, _lib.ctypes.byref(self.table), len(reqs), descs,
                                        stream.cuda_stream))
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev

    @staticmethod
    def _bucket_shape(sh):
        """The launch shape a graph is captured for: block counts rounded up to a
        power of two (at least 1024: blocks past a batch's own exit at once, so a
        small batch costs a few microseconds more on a larger grid), LDS bytes to a
        multiple of 8 KB.  Coarse buckets keep the graphs of a stream few."""
        def blocks(n):
            return 0 if n == 0 else max(1024, 1 << (int(n) - 1).bit_length())

        def lds(n):
            return (int(n) + 8191) // 8192 * 8192
        out = RRBatchShape(setup_blocks=blocks(sh.setup_blocks), setup_multi=sh.setup_multi,
                           setup_lds=lds(sh.setup_lds))
        for v in range(2):
            out.flux_blocks[v] = blocks(sh.flux_blocks[v])
            out.flux_lds[v] = lds(sh.flux_lds[v]) if sh.flux_blocks[v] else 0
        key = (out.setup_blocks, out.setup_multi, out.setup_lds, out.flux_blocks[0], out.flux_lds[0],
               out.flux_blocks[1], out.flux_lds[1])
        return key, out

    def _replay(self, slot, stream, reqs, descs):
        """Plan the batch into the slot's pinned descriptor, copy it to the slot's
        device descriptor, then replay (capturing first if new) the graph of the
        bucketed launch shape.  The pinned descriptor is rewritten only once the
        slot's previous batch has completed (the slot is free again only then)."""
        lib = self.lib
        prec = self.server._prec
        shape = RRBatchShape()
        check(lib.rr_batch_plan(prec, _lib.ctypes.byref(self.table), len(reqs), descs,
                                self.desc_host[slot].data_ptr(), _lib.ctypes.byref(shape)))
        with torch.cuda.stream(stream):
            self.desc_dev[slot].copy_(self.desc_host[slot], non_blocking=True)
        key, launch = self._bucket_shape(shape)
        graphs = self.slot_graphs[slot]
        g = graphs.get(key)
        if g is None:
            # captured on a stream of its own that runs nothing else, without the
            # device synchronisation and garbage collection of torch.cuda.graph():
            # the batches in flight on other slots are not waited for
            g = torch.cuda.CUDAGraph()
            desc = self.desc_dev[slot].data_ptr()
            with torch.cuda.stream(self.capture_stream):
                g.capture_begin(capture_error_mode='thread_local')
                try:
                    check(lib.rr_batch_launch(prec, desc, _lib.ctypes.byref(launch),
                                              self.capture_stream.cuda_stream))
                finally:
                    g.capture_end()
            graphs[key] = g
            self.graph_captures += 1
        with torch.cuda.stream(stream):
            g.replay()
        self.graph_replays += 1

    @staticmethod
    def finished(event) -> bool:
        return event.query()
