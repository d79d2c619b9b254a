"""Batched ensemble: many independent replicas of uneven size advanced in the
same kernel launches (one pool), versus the same replicas run one at a time.

Each launch of a step phase covers every replica's beads; per-replica step
state (dt, step count, max force) lives on the device, so replicas of any
size and parameters share launches without synchronising with each other.
"""
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
                   warmup=None):
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
