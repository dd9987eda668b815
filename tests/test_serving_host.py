"""Host logic of the request server: the arena allocator and the scheduler
(admission by memory budget, batching, slots, skip limit, dataset residency),
driven by a fake executor so no GPU is needed."""
import numpy as np
import pytest

from pytransit_amd.serving import ALIGN, Dataset, FreeList, RRServer


# ---------------------------------------------------------------------------------
def test_freelist_alignment_and_coalescing():
    fl = FreeList(10 * ALIGN + 17)
    assert fl.capacity == 10 * ALIGN
    a = fl.alloc(1)
    b = fl.alloc(ALIGN + 1)
    c = fl.alloc(3 * ALIGN)
    assert (a, b, c) == (0, ALIGN, 3 * ALIGN)
    assert fl.in_use == 6 * ALIGN
    assert fl.alloc(5 * ALIGN) is None
    fl.free(b)
    assert fl.alloc(2 * ALIGN) == ALIGN  # first fit reuses the hole
    fl.free(ALIGN)
    fl.free(a)
    fl.free(c)
    assert fl.in_use == 0 and fl._free == [(0, 10 * ALIGN)]


def test_freelist_random_against_interval_model():
    rng = np.random.default_rng(3)
    cap = 4096 * ALIGN
    fl = FreeList(cap)
    live = {}
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off = list(live)[rng.integers(len(live))]
            fl.free(off)
            del live[off]
        else:
            n = int(rng.integers(1, 64 * ALIGN))
            off = fl.alloc(n)
            if off is None:
                # first fit fails only if no free gap is large enough
                assert fl.largest_free() < -(-n // ALIGN) * ALIGN
                continue
            assert off % ALIGN == 0 and off + n <= cap
            for o, m in live.items():
                assert off + n <= o or o + m <= off, 'overlapping blocks'
            live[off] = n
        assert fl.in_use == sum(-(-m // ALIGN) * ALIGN for m in live.values())
        gaps = fl._free
        assert all(gaps[j][0] + gaps[j][1] < gaps[j + 1][0] for j in range(len(gaps) - 1)), 'uncoalesced'
    for off in list(live):
        fl.free(off)
    assert fl._free == [(0, fl.capacity)]


# ---------------------------------------------------------------------------------
class FakeEvent:
    def __init__(self):
        self.done = False


class FakeExec:
    """Records launches; a batch finishes when the test says so."""
    max_batch = 16

    def __init__(self, server_ref, staging_bytes=1 << 20, ws_per_pv=1024):
        self.server_ref = server_ref
        self.staging_bytes = staging_bytes
        self.ws_per_pv = ws_per_pv
        self.launches = []
        self.uploads = []
        self.retired = []

    def workspace_bytes(self, r):
        return self.ws_per_pv * r.npv

    def staging_need(self, r):
        return r.params.nbytes

    def upload_dataset(self, ds):
        self.uploads.append(ds)
        ds.ready = FakeEvent()

    def retire_dataset(self, ds):
        self.retired.append(ds)

    def launch(self, slot, reqs):
        srv = self.server_ref[0]
        assert srv._arena.in_use <= srv._arena.capacity
        assert len(reqs) <= srv.max_batch
        assert sum(1 for b in srv._batches) < srv.slots
        for r in reqs:
            assert r.dataset.offset is not None and r.offset is not None and r.flux_offset is not None
        ev = FakeEvent()
        self.launches.append((slot, list(reqs), ev))
        return ev

    @staticmethod
    def finished(ev):
        return ev.done


def make_server(budget, slots=2, max_batch=4, max_skips=3, **kw):
    ref = [None]
    ex = FakeExec(ref, **kw)
    srv = RRServer(hbm_budget=budget, slots=slots, max_batch=max_batch, precision='mixed', executor=ex,
                   max_skips=max_skips, auto_release=True)
    ref[0] = srv
    return srv, ex


def dataset(npt=1000, npb=1):
    t = 2459000.0 + np.arange(npt) * (2 / 1440)
    lcids = (np.arange(npt) * npb) // npt
    return Dataset(t, lcids=lcids, pbids=np.arange(npb), nsamples=1, exptimes=0.0)


def submit(srv, ds, npv, arrival=0.0):
    return srv.submit(ds, np.full(npv, 0.1), np.tile([0.4, 0.2], (npv, 1, 1)), np.full(npv, 2459000.1),
                      np.full(npv, 3.0), np.full(npv, 10.0), np.full(npv, 1.5), ld='quadratic', arrival=arrival)


def finish_all(ex):
    for _, _, ev in ex.launches:
        ev.done = True


def drain(srv, ex, max_steps=1000):
    for _ in range(max_steps):
        if srv.idle:
            return
        finish_all(ex)
        srv.step()
    raise AssertionError('server did not drain')


def test_batches_slots_and_fifo_order():
    srv, ex = make_server(1 << 30, slots=2, max_batch=4)
    ds = dataset()
    reqs = [submit(srv, ds, 16) for _ in range(11)]
    srv.step()
    # two slots, four requests each, in arrival order
    assert [[r.rid for r in b] for _, b, _ in ex.launches] == [[0, 1, 2, 3], [4, 5, 6, 7]]
    srv.step()
    assert len(ex.launches) == 2  # no free slot
    ex.launches[0][2].done = True
    srv.step()
    assert [r.rid for r in ex.launches[2][1]] == [8, 9, 10]
    assert ex.launches[2][0] == ex.launches[0][0]  # the freed slot
    drain(srv, ex)
    assert all(r.done for r in reqs) and len(srv.completed) == 11
    assert len(ex.uploads) == 1  # the dataset stayed resident
    assert srv._arena.in_use == ds.nbytes // ALIGN * ALIGN + (ALIGN if ds.nbytes % ALIGN else 0)


def test_budget_admission_and_skip_limit():
    ds = dataset(npt=1000)
    # flux of one npv=64 request: 64*1000*4 = 256000 B; budget fits about two
    srv, ex = make_server(700_000, slots=4, max_batch=16, max_skips=2)
    big = [submit(srv, ds, 64) for _ in range(3)]
    small = [submit(srv, ds, 4) for _ in range(6)]
    srv.step()
    first = [r.rid for r in ex.launches[0][1]]
    # the third big request does not fit; smaller ones behind it are admitted
    assert first[:2] == [0, 1] and 2 not in first
    assert srv._arena.in_use <= srv._arena.capacity
    assert big[2].skips >= 1
    # keep the memory busy: later small requests pass the big one until the limit
    for _ in range(6):
        submit(srv, ds, 4)
        srv.step()
    assert big[2].skips <= 2
    assert not big[2].done and big[2].t_admit is None
    queued_after = [r for r in srv._queue if r.rid > 2]
    assert queued_after, 'later requests are held back once the skip limit is reached'
    drain(srv, ex)
    assert big[2].done
    assert all(r.done for r in small)


def test_request_larger_than_budget_is_rejected():
    srv, _ = make_server(100_000)
    ds = dataset(npt=1000)
    with pytest.raises(ValueError, match='budget'):
        submit(srv, ds, 64)


def test_dataset_eviction_lru_only_when_idle():
    d1, d2, d3 = dataset(npt=20000), dataset(npt=20000), dataset(npt=20000)
    # each dataset is ~400 KB (times + lcids); the budget holds two plus requests
    srv, ex = make_server(2 * d1.nbytes + 200_000, slots=4, max_batch=1)
    r1 = submit(srv, d1, 1)
    srv.step()
    r2 = submit(srv, d2, 1)
    srv.step()
    r3 = submit(srv, d3, 1)
    srv.step()
    # both resident datasets are in use: d3 must wait
    assert r3.t_admit is None and d3.offset is None
    ex.launches[0][2].done = True  # r1 finishes: d1 becomes idle and is evicted
    srv.step()
    assert r3.t_admit is not None and d3.offset is not None and d1.offset is None and d2.offset is not None
    drain(srv, ex)
    assert [d.users for d in (d1, d2, d3)] == [0, 0, 0]
    # a new request on d1 uploads it again
    submit(srv, d1, 1)
    srv.step()
    assert ex.uploads.count(d1) == 2


def test_idle_datasets_are_evicted_for_request_memory():
    """Idle resident datasets give way to a request's parameters, workspace and flux,
    not only to another dataset; the victim's upload is retired (waited for) first."""
    d1, d2 = dataset(npt=20000), dataset(npt=20000)
    budget = 2 * d1.nbytes + 200_000
    srv, ex = make_server(budget, slots=4, max_batch=1, ws_per_pv=1024)
    for d in (d1, d2):
        submit(srv, d, 1)
        srv.step()
    finish_all(ex)
    srv.step()
    assert d1.offset is not None and d2.offset is not None and d1.users == d2.users == 0
    # 3 pv on d1: 3 * 20000 * 4 B of flux does not fit beside two resident datasets
    r = submit(srv, d1, 3)
    assert r.flux_bytes + d1.nbytes < budget
    srv.step()
    assert r.t_admit is not None, 'the idle dataset d2 must be evicted for the request'
    assert d2.offset is None and d1.offset is not None and ex.retired == [d2]
    drain(srv, ex)
    assert r.done


def test_run_raises_when_unreleased_results_hold_the_arena():
    ref = [None]
    ex = FakeExec(ref)
    srv = RRServer(hbm_budget=700_000, slots=2, max_batch=4, precision='mixed', executor=ex, auto_release=False)
    ref[0] = srv
    ds = dataset(npt=1000)
    a = submit(srv, ds, 64)
    srv.step()
    finish_all(ex)
    srv.step()
    assert a.done                                          # its flux is still held
    b = submit(srv, ds, 64)
    c = submit(srv, ds, 64)
    srv.step()
    finish_all(ex)
    with pytest.raises(RuntimeError, match='cannot be admitted'):
        srv.run()
    a.release()
    b.release()
    drain(srv, ex)
    assert c.done


def test_request_validation():
    srv, _ = make_server(1 << 30)
    ds = dataset(npb=3)
    with pytest.raises(ValueError, match='Radius ratios'):
        srv.submit(ds, np.full((4, 2), 0.1), np.tile([0.4, 0.2], (4, 3, 1)), np.zeros(4), np.full(4, 3.0),
                   np.full(4, 10.0), np.full(4, 1.5))
    with pytest.raises(ValueError, match='ldc'):
        srv.submit(ds, np.full((4, 3), 0.1), np.tile([0.4, 0.2], (4, 2, 1)), np.zeros(4), np.full(4, 3.0),
                   np.full(4, 10.0), np.full(4, 1.5))
    with pytest.raises(KeyError):
        srv.submit(ds, 0.1, [0.4, 0.2], 0.0, 3.0, 10.0, 1.5, ld='no-such-law')
    r = srv.submit(ds, np.full((4, 3), 0.1), [0.4, 0.2], np.zeros(4), np.full(4, 3.0), 10.0, 1.5,
                   ld='quadratic')
    assert r.host['ldc'].shape == (4, 3, 2) and r.host['a'].shape == (4,)


def test_batch_points_cap_runs_large_requests_alone():
    ref = [None]
    ex = FakeExec(ref)
    srv = RRServer(hbm_budget=1 << 30, slots=2, max_batch=8, precision='mixed', executor=ex, max_skips=3,
                   auto_release=True, max_batch_points=50_000)
    ref[0] = srv
    ds = dataset(npt=1000)
    small1 = submit(srv, ds, 10)     # 10 000 points
    big = submit(srv, ds, 200)       # 200 000 points: above the cap, alone
    small2 = submit(srv, ds, 20)     # 20 000 points
    small3 = submit(srv, ds, 30)     # 30 000 points: would pass the cap with small1 + small2
    srv.step()
    first, second = [[r.rid for r in reqs] for _, reqs, _ in ex.launches[:2]]
    # the small requests before the cap share a batch; the large one heads the next
    assert first == [small1.rid, small2.rid]
    assert second == [big.rid]
    drain(srv, ex)
    third = [r.rid for r in ex.launches[2][1]]
    assert third == [small3.rid]
    assert all(r.done for r in (small1, big, small2, small3))
    assert sum(r.points for r in (small1, small2)) <= 50_000
