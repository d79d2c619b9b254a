"""Batched ensemble: many independent replicas of uneven size advanced in the
same kernel launches (one pool), versus the same replicas run one at a time.

Each launch of a step phase covers every replica's beads; per-replica step
state (dt, step count, max force) lives on the device, so replicas of any
size and parameters share launches without synchronising with each other.
"""
import copy
import time

import numpy as np

from . import engine as E
from .workloads import Workload


def pool_for(ws, precision=E.GCR_MIXED, extra_replicas=0):
    """A pool sized for the given workloads (2D or 3D, not mixed)."""
    dim = ws[0].phys.dim
    assert all(w.phys.dim == dim for w in ws)
    from . import _lib
    import ctypes as C
    lib = _lib.load()
    cells = 0
    for w in ws:
        d = w.phys.desc()
        cells += int(lib.gcr_replica_cells(dim, C.byref(d))) + 1
    nb = sum(w.n for w in ws)
    nch = sum(w.phys.n_chains for w in ws)
    return E.Pool(dim, precision, bead_capacity=nb, cell_capacity=cells,
                  chain_capacity=nch, replica_capacity=len(ws) + extra_replicas)


def add_all(pool, ws):
    rids = []
    for w in ws:
        rids.append(pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths,
                             w.v_tan_angle, w.v0_per_chain))
    return rids


def run_batched(ws, steps, precision=E.GCR_MIXED, steps_per_batch=64, rebuild_every=8,
                warmup=0):
    """All replicas in one pool. Returns (seconds, bead-steps, pool, rids)."""
    pool = pool_for(ws, precision)
    pool.set_schedule(steps_per_batch, rebuild_every)
    rids = add_all(pool, ws)
    if warmup:
        for w, r in zip(ws, rids):
            tv = w.time_vals.copy()
            tv[7] = warmup
            pool.begin(r, tv)
        pool.run()
    for w, r in zip(ws, rids):
        tv = w.time_vals.copy()
        if warmup:
            tv[0] = pool.status(r)["dt"]
        tv[7] = steps
        pool.begin(r, tv)
    t0 = time.perf_counter()
    pool.run()
    sec = time.perf_counter() - t0
    for r in rids:
        assert pool.status(r)["steps"] == steps
    return sec, sum(w.n for w in ws) * steps, pool, rids


def run_sequential(ws, steps, precision=E.GCR_MIXED, steps_per_batch=64, rebuild_every=8,
                   warmup=None, tile=0, threads=0):
    """Each replica alone in its own pool, one after another (same kernels).
    Like run_batched, each pool first runs `warmup` untimed steps (default
    one graph batch), which captures its batch graphs, so that both sides
    time steps only."""
    if warmup is None:
        warmup = steps_per_batch
    total_sec, total_bs = 0.0, 0
    for w in ws:
        pool = pool_for([w], precision)
        try:
            pool.set_schedule(steps_per_batch, rebuild_every)
            pool.set_tile(tile)
            pool.set_tile_threads(threads)
            (r,) = add_all(pool, [w])
            tv = w.time_vals.copy()
            if warmup:
                tv[7] = warmup
                pool.begin(r, tv)
                pool.run()
                tv[0] = pool.status(r)["dt"]
            tv[7] = steps
            pool.begin(r, tv)
            t0 = time.perf_counter()
            pool.run()
            total_sec += time.perf_counter() - t0
            assert pool.status(r)["steps"] == steps
            total_bs += w.n * steps
        finally:
            pool.close()
    return total_sec, total_bs


def summary(ws):
    n = np.array([w.n for w in ws])
    return dict(replicas=len(ws), beads=int(n.sum()), min_beads=int(n.min()),
                max_beads=int(n.max()), mean_beads=float(n.mean()))


def tune_pool(pool, ws, rids, tiles=(0, 4, 6, 8), threads=(256, 512), rebuilds=(8, 16, 32, 64, 128),
              batch=64, warm=128, kernel_steps=16, schedule_steps=256):
    """Autotunes a batched pool in place: the tile edge and threads per tile
    of its 2D force kernels (by the force-stage durations of every replica
    together, gcr_pool_time_kernels, on the state `warm` steps in; no result
    depends on them), then the re-sort period (by the wall time of graph
    batches, replays included). Returns the settings and the trials; the
    replicas are left armed for nothing (begin them again to run)."""
    def arm(steps):
        for w, r in zip(ws, rids):
            tv = w.time_vals.copy()
            tv[4] = -1.0
            tv[7] = steps
            pool.begin(r, tv)

    pool.set_schedule(batch, min(8, batch))
    arm(warm)
    pool.run()
    tiling = []
    for t in (tiles if pool.dim == 2 else (0,)):
        for nt in (threads if pool.dim == 2 else (0,)):
            pool.set_tile(t)
            pool.set_tile_threads(nt)
            arm(1 << 30)
            pool.time_kernels(2)  # tables, warm-up
            kt = pool.time_kernels(kernel_steps)
            tiling.append((t, nt, kt["stage0"] + kt["stage1"]))
    tiling.sort(key=lambda x: x[2])
    tile, nt = tiling[0][0], tiling[0][1]
    pool.set_tile(tile)
    pool.set_tile_threads(nt)
    sched = []
    for k in rebuilds:
        # a batch starts with a re-sort: a longer period runs in batches of
        # the period
        kb = max(batch, k)
        pool.set_schedule(kb, k)
        arm(kb)
        pool.run()  # captures
        n = max(schedule_steps, 2 * kb)
        arm(n)
        t0 = time.perf_counter()
        pool.run()
        sched.append((k, 1e3 * (time.perf_counter() - t0) / n))
    sched.sort(key=lambda x: x[1])
    best = sched[0][0]
    pool.set_schedule(max(batch, best), best)
    return dict(tile=tile, threads=nt, rebuild_every=best, steps_per_batch=max(batch, best),
                tiling_trials=[[t, n, round(1e3 * m, 2)] for t, n, m in tiling],
                schedule_trials=[[k, round(m, 4)] for k, m in sched])


def with_skin(ws, frac):
    """Copies of the workloads whose cell-list skin is `frac` times each
    replica's pair cutoff (positions and the rest shared)."""
    out = []
    for w in ws:
        c = copy.copy(w)
        c.phys = copy.copy(w.phys)
        c.phys.skin = frac * float(w.phys.wca[3])
        out.append(c)
    return out


def tune_skin(ws, precision=E.GCR_MIXED, fracs=(0.3, 0.6, 1.0, 1.4), steps=512, batch=64, **kw):
    """The cell-list skin of a batched pool, as a fraction of each replica's
    pair cutoff: for each candidate a pool of all replicas with that skin,
    tuned by tune_pool, then timed over `steps` steps of graph batches
    (replays included). The skin sets the cell grid, so each candidate is a
    pool of its own; results move at the rounding level only (another grid
    sums the same pairs in another order). Returns (best fraction, the
    tune_pool settings there, trials [fraction, ms per step, replays])."""
    trials, best = [], None
    for f in fracs:
        wf = with_skin(ws, f)
        pool = pool_for(wf, precision)
        try:
            rids = add_all(pool, wf)
            tuned = tune_pool(pool, wf, rids, batch=batch, **kw)
            kb = tuned["steps_per_batch"]
            for w, r in zip(wf, rids):
                tv = w.time_vals.copy()
                tv[4] = -1.0
                tv[7] = kb
                pool.begin(r, tv)
            pool.run()  # the tuned schedule's graphs
            replays0 = sum(pool.status(r)["replays"] for r in rids)
            for w, r in zip(wf, rids):
                tv = w.time_vals.copy()
                tv[4] = -1.0
                tv[7] = steps
                pool.begin(r, tv)
            t0 = time.perf_counter()
            pool.run()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            replays = sum(pool.status(r)["replays"] for r in rids) - replays0
        finally:
            pool.close()
        trials.append([f, round(ms, 4), int(replays)])
        if best is None or ms < best[1]:
            best = (f, ms, tuned)
    return best[0], best[2], trials
