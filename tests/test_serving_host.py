"""The serving loop's host logic (gcr_amd.serving.Server.serve) on the CPU:
admission under a pool's capacity, retirement, failed requests, errors, and
each pool advancing at its own pace. The pools are stand-ins that count
steps instead of running kernels; the GPU tests cover the real pools."""
import threading
import time
from types import SimpleNamespace

import pytest

from gcr_amd import serving as SV

STEPS_PER_BATCH = 64


class FakePool:
    def __init__(self, lane, delay, fail_at):
        self.lane, self.delay, self.fail_at = lane, delay, fail_at

    def run(self, max_batches=0):
        assert max_batches >= 1
        self.calls += 1
        for b in range(max_batches):
            if not any(v > 0 for v in self.lane.left.values()):
                break
            time.sleep(self.delay)
            if self.fail_at is not None and self.done == self.fail_at:
                raise RuntimeError(f"pool of dim {self.lane.dim} failed")
            self.done += 1
            for rid in self.lane.active:
                self.lane.left[rid] -= STEPS_PER_BATCH


class FakeLane(SV._Lane):
    """The interface Server.serve drives (fits, admit, retire, pool.run, and
    _Lane's batches_ahead)."""

    def __init__(self, dim, cap_beads, delay=0.0, fail_at=None, bad_n=None):
        self.dim = dim
        self.cap_beads = cap_beads
        self.steps_per_batch = STEPS_PER_BATCH
        self.used = 0
        self.active = {}
        self.left = {}
        self.rejected = []
        self.batches = 0
        self.busy_s = 0.0
        self.batch_s = 0.0
        self.min_left = None
        self.bad_n = bad_n
        self.peak = 0
        self.pool = FakePool(self, delay, fail_at)
        self.pool.calls = self.pool.done = 0
        self.thread = set()

    def fits(self, req, cells):
        return self.used + req.n <= self.cap_beads

    def admit(self, req, cells, now):
        self.thread.add(threading.current_thread())
        if req.n == self.bad_n:  # the library refused the input
            req.admitted = req.finished = now
            req.result = {"status": {"failed": True, "steps": 0, "error": "bad input"}}
            self.rejected.append(req)
            return True
        self.active[req.rid] = (req, cells)
        self.left[req.rid] = req.steps
        self.used += req.n
        self.peak = max(self.peak, self.used)
        if self.min_left is not None:
            self.min_left = min(self.min_left, req.steps)
        req.admitted = now
        return True

    def retire(self, now, keep_results):
        done = []
        for rid, (req, _) in list(self.active.items()):
            if self.left[rid] <= 0:
                req.finished = now
                req.result = {"status": {"failed": False, "steps": req.steps}}
                self.used -= req.n
                del self.active[rid]
                del self.left[rid]
                done.append(req)
        self.min_left = min(self.left.values()) if self.left else None
        return done


def request(rid, dim, n, steps, arrival=0.0):
    work = SimpleNamespace(n=n, phys=SimpleNamespace(dim=dim, n_chains=max(1, n // 10)))
    return SV.Request(rid, arrival, work, steps)


def server(lanes):
    srv = SV.Server.__new__(SV.Server)
    srv.skin_frac = None
    srv.lanes = lanes
    return srv


@pytest.fixture(autouse=True)
def no_library(monkeypatch):
    monkeypatch.setattr(SV, "replica_cells", lambda w: 1)


def test_every_request_finishes_once_under_capacity():
    lanes = {2: FakeLane(2, cap_beads=3000), 3: FakeLane(3, cap_beads=2500)}
    reqs = [request(i, 2 if i % 2 == 0 else 3, 400 + 150 * (i % 7), 64 * (1 + i % 5))
            for i in range(30)]
    done = server(lanes).serve(reqs, realtime=False)
    assert sorted(r.rid for r in done) == list(range(30))
    for r in done:
        assert not r.result["status"]["failed"]
        assert 0.0 <= r.admitted <= r.finished
    # the capacity made requests wait for room, and was never exceeded
    assert lanes[2].peak <= 3000 and lanes[3].peak <= 2500
    assert max(r.admitted for r in done) > 0.0
    # one host thread per pool, and not the caller's
    assert len(lanes[2].thread) == 1 and len(lanes[3].thread) == 1
    assert lanes[2].thread.isdisjoint(lanes[3].thread)
    assert threading.current_thread() not in lanes[2].thread | lanes[3].thread


def test_pools_advance_at_their_own_pace():
    # a slow 3D pool (50 ms per batch) does not hold back the 2D pool (1 ms)
    lanes = {2: FakeLane(2, 10**6, delay=0.001), 3: FakeLane(3, 10**6, delay=0.05)}
    reqs = [request(0, 3, 1000, 64 * 10)] + [request(1 + i, 2, 500, 64 * 40) for i in range(3)]
    done = server(lanes).serve(reqs, realtime=False)
    fin = {r.rid: r.finished for r in done}
    assert lanes[2].pool.done == 40 and lanes[3].pool.done == 10
    # 40 batches of the 2D pool at 1 ms end well before 10 of the 3D pool at 50 ms;
    # stepped together they would take 40 x 50 ms
    assert max(fin[i] for i in (1, 2, 3)) < 0.5 * fin[0]
    assert max(fin.values()) < 1.5


def test_rejected_request_is_reported_failed():
    lanes = {2: FakeLane(2, 10**6, bad_n=777), 3: FakeLane(3, 10**6)}
    reqs = [request(0, 2, 500, 64), request(1, 2, 777, 64), request(2, 3, 300, 128)]
    done = server(lanes).serve(reqs, realtime=False)
    st = {r.rid: r.result["status"] for r in done}
    assert sorted(st) == [0, 1, 2]
    assert st[1]["failed"] and "bad input" in st[1]["error"]
    assert not st[0]["failed"] and not st[2]["failed"]


def test_error_in_one_pool_ends_the_stream():
    lanes = {2: FakeLane(2, 10**6, delay=0.001), 3: FakeLane(3, 10**6, fail_at=2)}
    reqs = [request(0, 2, 500, 64 * 1000), request(1, 3, 500, 64 * 10)]
    t0 = time.perf_counter()
    with pytest.raises(RuntimeError, match="pool of dim 3 failed"):
        server(lanes).serve(reqs, realtime=False)
    # the 2D pool's thread stops at its next return to the loop instead of
    # running its 1000 batches
    assert lanes[2].pool.done < 200
    assert time.perf_counter() - t0 < 5


def test_request_beyond_a_pool_is_refused_up_front():
    lanes = {2: FakeLane(2, 1000), 3: FakeLane(3, 1000)}
    with pytest.raises(ValueError, match="exceeds the HBM budget"):
        server(lanes).serve([request(0, 3, 5000, 64)], realtime=False)


def test_arrivals_are_respected_in_realtime():
    lanes = {2: FakeLane(2, 10**6), 3: FakeLane(3, 10**6)}
    reqs = [request(0, 2, 100, 64, arrival=0.0), request(1, 2, 100, 64, arrival=0.15),
            request(2, 3, 100, 64, arrival=0.1)]
    done = server(lanes).serve(reqs, realtime=True)
    for r in done:
        assert r.admitted >= r.arrival - 1e-3
    assert {r.rid for r in done} == {0, 1, 2}


def test_several_batches_per_run_until_the_next_completion():
    # requests of 3, 5 and 8 batches admitted together: the loop runs 3
    # batches, retires one, runs 2, retires one, runs 3 -- no return to the
    # host between batches that complete nothing
    lanes = {2: FakeLane(2, 10**6), 3: FakeLane(3, 10**6)}
    reqs = [request(0, 2, 100, 64 * 3), request(1, 2, 100, 64 * 5 - 10), request(2, 2, 100, 64 * 8)]
    done = server(lanes).serve(reqs, realtime=False)
    assert sorted(r.rid for r in done) == [0, 1, 2]
    assert lanes[2].pool.done == 8
    # the first call runs one batch (nothing known yet), then 2, 2 and 3
    assert lanes[2].pool.calls == 4


def test_batches_ahead_stops_at_the_next_arrival():
    lane = FakeLane(2, 10**6)
    lane.min_left = 64 * 100
    assert lane.batches_ahead([], 0.0, realtime=True) == SV.MAX_BATCHES_PER_RUN
    assert lane.batches_ahead([request(0, 2, 10, 64, arrival=5.0)], 0.0, realtime=True) == 1
    lane.batch_s = 0.1
    assert lane.batches_ahead([request(0, 2, 10, 64, arrival=1.05)], 0.0, realtime=True) == 10
    assert lane.batches_ahead([request(0, 2, 10, 64, arrival=1.05)], 0.0, realtime=False) == 32
    lane.min_left = 64 * 3 - 1
    assert lane.batches_ahead([], 0.0, realtime=True) == 3
    lane.min_left = None
    assert lane.batches_ahead([], 0.0, realtime=True) == 1
