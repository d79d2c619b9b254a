"""Autotuner of the cell list and the 2D tiles for one workload.

Picks by timing on the device:

- the tile edge and threads per tile of the 2D force kernels, from the
  force-stage durations (gcr_pool_time_kernels). No result bit depends on
  them: every pair path forms separations from cell-relative positions and
  sums in cell-scan order, whatever the tiling;
- the cell-list skin (cell edge >= cutoff + skin; the neighbour-list radius is
  the cell edge) and the re-sort period, from whole-step wall time of graph
  batches, replays included (a skin too thin for the period makes batches
  replay with a re-sort before every step). These move results at the
  rounding level only: another cell grid sums the same pairs in another order.

Every candidate runs on a temporary pool holding a copy of the workload's
initial state; no pool of the caller is touched. The returned settings go to
Physics.skin, Pool.set_schedule, Pool.set_tile and Pool.set_tile_threads.
"""
import copy
import time

from . import engine as E
from ._lib import GcrError


def _pool_for(w, precision, skin):
    phys = copy.copy(w.phys)
    phys.skin = skin
    n = w.n
    cells = max(4 * n + 1024, E.replica_cells(phys) + 1)
    pool = E.Pool(phys.dim, precision, bead_capacity=n, cell_capacity=cells,
                  chain_capacity=phys.n_chains, replica_capacity=1)
    rid = pool.add(phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                   getattr(w, "v0_per_chain", None))
    return pool, rid


def _arm(pool, rid, w, steps):
    tv = w.time_vals.copy()
    tv[4] = -1.0  # no time limit: exactly `steps` steps
    tv[7] = steps
    pool.begin(rid, tv)


def tune_tiling(w, precision, skin, tiles=(0, 4, 6, 8, 10, 12), threads=(256, 512, 1024), steps=32,
                warm=256):
    """[(tile, threads, stage0 + stage1 ms)] for every candidate that ran
    without exceeding the skin, fastest first. The candidates are timed on
    the state `warm` steps after the initial one (chains of attractive beads
    cluster within the first hundreds of steps, which lengthens the lists of
    dense tiles); each timing restores that state, so every candidate sees
    the same one. In 2D, after the squares, rectangles of about the best
    square's area and more cells in x (tile = (x, y)) at its thread count:
    a halo row of more cells is a longer run of slots to load."""
    pool, rid = _pool_for(w, precision, skin)
    out = []

    def trial(t, nt):
        pool.set_tile(t)
        pool.set_tile_threads(nt)
        _arm(pool, rid, w, 1 << 30)
        try:
            pool.time_kernels(4)  # builds tables and warms up
            kt = pool.time_kernels(steps)
        except GcrError:
            return
        out.append((t, nt, kt["stage0"] + kt["stage1"]))

    try:
        if warm > 0:
            pool.set_schedule(64, 8)
            _arm(pool, rid, w, warm)
            pool.run()
        for t in tiles:
            for nt in threads:
                trial(t, nt)
        best = min(out, key=lambda r: r[2]) if out else None
        if w.phys.dim == 2 and best is not None and isinstance(best[0], int) and best[0] > 0:
            for shape in rectangles(best[0]):
                trial(shape, best[1])
    finally:
        pool.close()
    return sorted(out, key=lambda r: r[2])


def rectangles(s, max_edge=16):
    """(x, y) tiles of about s x s cells with 1 to 3 more cells in x, or
    twice as many: y = floor or round of s^2 / x."""
    shapes = []
    for x in (s + 1, s + 2, s + 3, 2 * s):
        for y in (int(s * s / x), int(round(s * s / x))):
            if 1 <= y < x <= max_edge and (x, y) not in shapes:
                shapes.append((x, y))
    return shapes


def time_schedule(w, precision, skin, rebuild_every, tile, threads, steps=256, batch=64):
    """Wall-clock ms per step of graph batches (replays included) and the
    number of replays, after one untimed batch (graph capture). A batch
    starts with a re-sort, so a period longer than `batch` runs in batches
    of the period."""
    batch = max(batch, rebuild_every)
    steps = max(steps, 2 * batch)
    pool, rid = _pool_for(w, precision, skin)
    try:
        pool.set_schedule(batch, rebuild_every)
        pool.set_tile(tile)
        pool.set_tile_threads(threads)
        _arm(pool, rid, w, batch)
        pool.run()
        _arm(pool, rid, w, steps)
        t0 = time.perf_counter()
        pool.run()
        sec = time.perf_counter() - t0
        st = pool.status(rid)
    finally:
        pool.close()
    return 1e3 * sec / steps, st["replays"], st["failed"]


def tune(w, precision=E.GCR_MIXED, skin_fracs=(0.1, 0.15, 0.2, 0.3, 0.4), rebuilds=(8, 16, 32, 64, 128),
         tiles=(0, 4, 6, 8, 10, 12), threads=(256, 512, 1024), steps=256, tiles3=(0, 2, 3, 4),
         threads3=(256, 512)):
    """Best (skin, rebuild_every, tile, threads) for workload w.

    The tiling is tuned first at the default skin (0.3 cutoff), then every
    (skin, period) pair is timed with that tiling, and the tiling is re-tuned
    at the chosen skin. skin_fracs are fractions of the pair cutoff; tiles
    and threads are the 2D candidates, tiles3 and threads3 the 3D ones (cube
    sides; 0 is the automatic choice). A re-sort period longer than the
    caller's graph batch needs batches of that period (a batch starts with a
    re-sort)."""
    rc = float(w.phys.wca[3])
    if w.phys.dim == 3:
        tiles, threads = tiles3, threads3
    tiling = tune_tiling(w, precision, 0.3 * rc, tiles, threads)
    tile, nt = tiling[0][0], tiling[0][1]
    trials = []
    for f in skin_fracs:
        for k in rebuilds:
            ms, replays, failed = time_schedule(w, precision, f * rc, k, tile, nt, steps)
            if not failed:
                trials.append(dict(skin=f * rc, rebuild_every=k, ms_per_step=ms, replays=replays))
    best = min(trials, key=lambda r: r["ms_per_step"])
    tiling = tune_tiling(w, precision, best["skin"], tiles, threads)
    tile, nt = tiling[0][0], tiling[0][1]
    return dict(skin=best["skin"], rebuild_every=best["rebuild_every"], tile=tile, threads=nt,
                ms_per_step=best["ms_per_step"], schedule_trials=trials,
                tiling_trials=[dict(tile=t, threads=n, stages_ms=m) for t, n, m in tiling])
