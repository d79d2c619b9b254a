"""Serving: a request loop with continuous batching of simulations.

Requests (a simulation: initial state, physics, number of steps) arrive over
time. The server admits a request as soon as its pool has room for it under
the HBM budget (beads, cells, chains, replica slots reserved at start-up),
keeps its state resident on the device, advances every admitted simulation
in the same launches batch after batch, and retires a simulation the batch
it completes, freeing its room for the next arrivals. 2D and 3D requests go
to two pools, each driven by its own host thread at its own pace (admit,
one graph batch, retire), so their launches overlap on the device and
neither pool waits for the other's batches. Results do not depend on what else is running: noise is keyed by
(request seed, bead, step), reductions are per replica and order-free, and
the cell-list schedule follows each replica's own step count.
"""
import ctypes as C
import copy
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import engine as E
from . import workloads as W


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the stream starts
    work: W.Workload    # initial state + physics (the client's payload)
    steps: int
    admitted: float = -1.0
    finished: float = -1.0
    result: dict = field(default_factory=dict)

    @property
    def dim(self):
        return self.work.phys.dim

    @property
    def n(self):
        return self.work.n


def make_request_workload(rng, dim, n_target, abp, seed, T=0.01, phi=None):
    """ABPs (chain length 1, D_r = 0.1) or chiral chains (length 5..40).
    Default packing: area fraction 0.25 in 2D; volume fraction 0.25 for 3D
    ABPs and 0.15 for 3D chains (random placement of long rods jams early)."""
    if phi is None:
        phi = 0.25 if (dim == 2 or abp) else 0.15
    if abp:
        ln, nc = 1, max(1, n_target)
    else:
        ln = int(rng.integers(5, 41))
        nc = max(1, n_target // ln)
    vt = float(rng.uniform(-0.5, 0.5)) if not abp else 0.0
    if dim == 2:
        L = W.box_from_area_fraction(nc, ln, 0.5, 1.0, phi)
    else:  # volume fraction of beads of diameter 1
        L = (nc * ln * (np.pi / 6) / phi) ** (1.0 / 3.0)
    return W.chains(f"req{seed}", dim, nc, ln, L=L, vtan=vt, T=T, Dr=0.1 if abp else None,
                    dtmax=1e-3, seed=seed, min_dist_floor=0.7)


def make_stream(seed=0, n_requests=64, rate=20.0, n_min=1000, n_max=200000, steps_min=1000,
                steps_max=100000, scale_steps=1.0):
    """Poisson arrivals; bead counts and step counts log-uniform; half ABPs,
    half chains; half 2D, half 3D. scale_steps shortens every request."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        steps = int(np.exp(rng.uniform(np.log(steps_min), np.log(steps_max))) * scale_steps)
        dim = 2 if rng.random() < 0.5 else 3
        abp = rng.random() < 0.5
        w = make_request_workload(rng, dim, n, abp, seed=seed * 1000003 + i)
        reqs.append(Request(i, t, w, max(1, steps)))
    return reqs


def replica_cells(w):
    lib = _lib.load()
    d = w.phys.desc()
    return int(lib.gcr_replica_cells(w.phys.dim, C.byref(d))) + 1


# most graph batches one call of the serving loop runs before it looks at
# arrivals, errors and completions again
MAX_BATCHES_PER_RUN = 32


class _Lane:
    """One pool (one dimensionality) and its bookkeeping."""

    def __init__(self, dim, budget_bytes, precision, steps_per_batch, rebuild_every,
                 max_replicas):
        # split the byte budget: per bead ~ gcr_replica_bytes(dim, 1, 1, 2)
        lib = _lib.load()
        per_bead = int(lib.gcr_replica_bytes(dim, 1, 1, 2))
        self.cap_beads = max(1024, budget_bytes // per_bead)
        self.cap_cells = 2 * self.cap_beads + 1024
        self.cap_chains = self.cap_beads
        self.cap_reps = max_replicas
        self.pool = E.Pool(dim, precision, self.cap_beads, self.cap_cells, self.cap_chains,
                           max_replicas)
        self.pool.set_schedule(steps_per_batch, rebuild_every)
        self.steps_per_batch = steps_per_batch
        self.used = np.zeros(3, dtype=np.int64)  # beads, cells, chains
        self.active = {}  # rid -> (request, cells)
        self.batches = 0
        self.busy_s = 0.0  # host time inside the pool's graph batches
        self.batch_s = 0.0  # host time of the last graph batch
        self.min_left = None  # steps left of the active request nearest its end (None: unknown)
        self.rejected = []  # requests whose input the library refused

    def fits(self, req, cells):
        need = np.array([req.n, cells, req.work.phys.n_chains])
        cap = np.array([self.cap_beads, self.cap_cells, self.cap_chains])
        return len(self.active) < self.cap_reps and np.all(self.used + need <= cap)

    def admit(self, req, cells, now):
        """True: admitted; False: no contiguous room (fragmentation), retry
        later. A request the library rejects for another reason (its input)
        is finished at once as failed, with the message in its result."""
        w = req.work
        try:
            rid = self.pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths,
                                w.v_tan_angle, w.v0_per_chain)
        except _lib.GcrError as e:
            if e.rc == _lib.GCR_ERR_CAPACITY:
                return False
            req.admitted = req.finished = now
            req.result = {"status": {"failed": True, "steps": 0, "error": str(e)}}
            self.rejected.append(req)
            return True
        tv = w.time_vals.copy()
        tv[7] = req.steps
        self.pool.begin(rid, tv)
        self.used += [req.n, cells, w.phys.n_chains]
        self.active[rid] = (req, cells)
        if self.min_left is not None:
            self.min_left = min(self.min_left, req.steps)
        req.admitted = now
        return True

    def retire(self, now, keep_results):
        """Retires the requests that completed (or failed); records the steps
        left of the nearest-to-complete request still running."""
        done = []
        left = []
        for rid, (req, cells) in list(self.active.items()):
            st = self.pool.status(rid)
            if st["running"]:
                left.append(req.steps - st["steps"])
            else:
                if keep_results and not st["failed"]:
                    req.result = self.pool.download(rid, pos1=False, forces=False)
                req.result["status"] = st
                req.finished = now
                self.pool.remove(rid)
                self.used -= [req.n, cells, req.work.phys.n_chains]
                del self.active[rid]
                done.append(req)
        self.min_left = min(left) if left else None
        return done

    def batches_ahead(self, queue, now, realtime):
        """Graph batches the next run may take without delaying a retirement
        or an admission: up to the batch in which the nearest-to-complete
        request ends, and, while requests are still to arrive, not past the
        next arrival (at the last batch's duration); one when either is
        unknown. Fewer returns to the host between batches leave fewer gaps
        in the pool's launches."""
        if self.min_left is None or self.min_left <= 0:
            return 1
        k = -(-self.min_left // self.steps_per_batch)
        if queue and realtime:
            if self.batch_s <= 0.0:
                return 1
            k = min(k, int((queue[0].arrival - now) / self.batch_s))
        return max(1, min(k, MAX_BATCHES_PER_RUN))


class Server:
    """skin_frac: the cell-list skin the server gives every request, as a
    fraction of the request's pair cutoff (None: the request's own). The
    skin is an engine setting: results move at the rounding level only."""

    def __init__(self, hbm_budget_bytes=8 << 30, precision=E.GCR_MIXED, steps_per_batch=64,
                 rebuild_every=8, max_replicas=256, split_2d=0.5, skin_frac=None):
        self.skin_frac = skin_frac
        self.lanes = {
            2: _Lane(2, int(hbm_budget_bytes * split_2d), precision, steps_per_batch,
                     rebuild_every, max_replicas),
            3: _Lane(3, int(hbm_budget_bytes * (1 - split_2d)), precision, steps_per_batch,
                     rebuild_every, max_replicas),
        }

    def close(self):
        for lane in self.lanes.values():
            lane.pool.close()

    def serve(self, requests, keep_results=False, realtime=True):
        """Run the stream. With realtime, requests become visible at their
        arrival time; otherwise all are visible at once (offline batch)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        if self.skin_frac is not None:
            for r in pending:
                w = copy.copy(r.work)
                w.phys = copy.copy(r.work.phys)
                w.phys.skin = self.skin_frac * float(w.phys.wca[3])
                r.work = w
        for r in pending:
            if not self.lanes[r.dim].fits(r, 0) and r.n > self.lanes[r.dim].cap_beads:
                raise ValueError(f"request {r.rid} exceeds the HBM budget of its pool")
        cells = {r.rid: replica_cells(r.work) for r in pending}
        finished = []
        errors = []
        lock = threading.Lock()
        stop = threading.Event()  # set when a pool fails: the others end too
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def drive(lane, queue):
            """One lane's loop: admit what has arrived and fits, advance the
            lane's pool by as many graph batches as it can take before the
            next completion or arrival (batches_ahead), retire what
            completed. Each lane keeps its own pace: a batch of the 2D pool
            never waits for one of the 3D pool, and their launches overlap on
            the device."""
            waiting = []
            while (queue or waiting or lane.active) and not stop.is_set():
                t = now()
                while queue and (not realtime or queue[0].arrival <= t):
                    waiting.append(queue.pop(0))
                # admission in arrival order; a request that does not fit
                # waits, smaller later ones may pass it (backfill)
                still = []
                for r in waiting:
                    if lane.fits(r, cells[r.rid]) and lane.admit(r, cells[r.rid], now()):
                        continue
                    still.append(r)
                waiting = still
                done, lane.rejected = lane.rejected, []
                if lane.active:
                    k = lane.batches_ahead(queue, now(), realtime)
                    tb = time.perf_counter()
                    lane.pool.run(max_batches=k)
                    run_s = time.perf_counter() - tb
                    lane.busy_s += run_s
                    lane.batch_s = run_s / k
                    lane.batches += k
                    done += lane.retire(now(), keep_results)
                elif waiting:  # nothing runs and nothing fits: it never will
                    raise RuntimeError(f"requests {[r.rid for r in waiting]} cannot be admitted "
                                       "into an idle server")
                elif queue and realtime:
                    time.sleep(max(0.0, min(0.01, queue[0].arrival - now())))
                with lock:
                    finished.extend(done)

        def guarded(lane, queue):
            try:
                drive(lane, queue)
            except BaseException as e:  # re-raised by serve
                with lock:
                    errors.append(e)
                stop.set()

        threads = []
        for dim, lane in self.lanes.items():
            queue = [r for r in pending if r.dim == dim]
            if queue:
                threads.append(threading.Thread(target=guarded, args=(lane, queue)))
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        return finished


def report(finished):
    failed = [r for r in finished if r.result["status"]["failed"]]
    ok = [r for r in finished if not r.result["status"]["failed"]]
    lat = np.array([r.finished - r.arrival for r in ok])
    work = sum(r.n * r.result["status"]["steps"] for r in finished)
    span = max(r.finished for r in finished) - min(r.arrival for r in finished)
    return dict(requests=len(finished), failed=len(failed), bead_steps=int(work), makespan_s=span,
                bead_steps_per_s=work / span, p50_latency_s=float(np.percentile(lat, 50)),
                p99_latency_s=float(np.percentile(lat, 99)),
                mean_queue_s=float(np.mean([r.admitted - r.arrival for r in ok])))
