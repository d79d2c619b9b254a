#!/usr/bin/env python3
"""Benchmark: bead-steps/s of the SRK2 Brownian-dynamics step on one MI355X.

Workload: 2D periodic box, 20k chains x 20 beads (400k beads), area fraction
0.4, LJ with attraction (eps=1, cutoff 2.5 sigma), bending 5, chirality
U[-0.5, 0.5] per chain, speed 1, T=0.05, dt=5e-4. One step = one SRK2 step
(predictor + corrector) of every bead. `value` is the mixed form (FP32 pair
forces, FP64 accumulation and state); the plain FP64 form is timed beside it.
The same line carries the ensemble (256 uneven replicas in one pool, 10^4
steps each, against replicas run one at a time) and the serving stream (48
Poisson requests through the continuous-batching server, aggregate rate and
p50/p99 completion latency), and the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-ensemble] [--no-serving]

Prints one JSON line. Inputs are resident in HBM before every timed region.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gliding-collective-rotation_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per bead per launch (FP64 state, 2D): stage 0 reads x
# and writes x1, stage 1 reads x1, the update reads and writes x -> 80 B per
# bead-step in total
ALG_BYTES = {"stage0": 32, "stage1": 16, "update": 32}
# HBM bytes per bead per launch of the mixed kernels on this workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes (16 x 4-cell tiles of 256 threads,
# the autotuner's pick, skin 0.3 cutoff, re-sort every 32 steps), each
# counter scaled by its reading for a plain 1 GiB copy kernel (`--mode copy`:
# reads x 2.0, writes x 1.0); profiles/stage_kernels_400k_beads.md holds the
# passes and commands. Measured again whenever these kernels change.
TRAFFIC_BYTES = {"stage0": 208.0, "stage1": 144.0, "update": 105.0}


def parse_tile(text):
    """'8' -> 8, '16x4' -> (16, 4)."""
    if "x" in text:
        x, y = text.split("x")
        return (int(x), int(y))
    return int(text)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1024)
    ap.add_argument("--warmup", type=int, default=128)
    ap.add_argument("--batch", type=int, default=64, help="steps per captured graph")
    ap.add_argument("--rebuild", type=int, default=8, help="cell-list rebuild period")
    ap.add_argument("--skin", type=float, default=0.0)
    ap.add_argument("--tile", default="0",
                    help="tile edge in cells, or XxY for a 2D rectangle (0: automatic)")
    ap.add_argument("--threads", type=int, default=0, help="threads per tile (0: automatic)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="single: take skin / rebuild / tile / threads from the flags instead of "
                         "timing candidates first")
    ap.add_argument("--chains", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU-baseline budget: steps are added in chunks of 40 until this much time")
    ap.add_argument("--kernel-steps", type=int, default=32)
    ap.add_argument("--mode", choices=("single", "ensemble", "serve", "request", "copy",
                                       "features", "observables"),
                    default="single",
                    help="single: 400k-bead system (the headline line); ensemble: 256 "
                         "replicas batched vs one at a time; serve: request stream; request: "
                         "requests of the serving stream (--ids) each alone in a pool; copy: "
                         "a plain 1 GiB device copy kernel, the reference point for HBM "
                         "byte counters (FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--ids", default="4,29",
                    help="request mode: indices of the serving stream's requests to time")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=10000,
                    help="ensemble: steps of every replica (the configuration's 10^4)")
    ap.add_argument("--seq-sample", type=int, default=32,
                    help="ensemble: replicas timed one at a time for the sequential rate")
    ap.add_argument("--seq-steps", type=int, default=512,
                    help="ensemble: steps of each replica timed one at a time")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--steps-scale", type=float, default=1.0,
                    help="serve: multiplies every request's step count")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="single: leave the ensemble part out of the line")
    ap.add_argument("--no-serving", action="store_true",
                    help="single: leave the serving part out of the line")
    ap.add_argument("--budget-gb", type=float, default=16.0)
    ap.add_argument("--rate", type=float, default=20.0,
                    help="serving: mean request arrivals per second (Poisson)")
    ap.add_argument("--serve-skin", type=float, default=0.6,
                    help="serve: cell-list skin of every request as a fraction of its cutoff "
                         "(0: the requests' own, 0.3)")
    ap.add_argument("--serve-rebuild", type=int, default=32,
                    help="serving: steps between cell re-sorts of the server's pools")
    ap.add_argument("--only", default="",
                    help="features mode: run only the options whose name contains this")
    return ap.parse_args()


def run_gpu(w, precision, steps, warmup, batch, rebuild, kernel_steps, tile=0, threads=0):
    from gcr_amd import engine as E
    n = w.n
    pool = E.Pool(2, precision, bead_capacity=n, cell_capacity=4 * n + 1024,
                  chain_capacity=w.phys.n_chains, replica_capacity=1)
    # a batch starts with a re-sort: a longer re-sort period needs batches of
    # that period; a run shorter than one graph batch gets a batch of its own
    # length, so the timed region holds no launches of steps that do not run
    batch = min(max(batch, rebuild), steps)
    rebuild = min(rebuild, batch)
    pool.set_schedule(batch, rebuild)
    pool.set_tile(tile)
    pool.set_tile_threads(threads)
    rid = pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
    tv = w.time_vals.copy()
    tv[7] = max(warmup, batch)  # captures the batch graphs before the clock starts
    pool.begin(rid, tv)
    pool.run()
    st = pool.status(rid)
    tv[0] = st["dt"]
    tv[7] = steps
    pool.begin(rid, tv)
    t0 = time.perf_counter()
    pool.run()  # returns after the device finished the last batch
    t1 = time.perf_counter()
    st2 = pool.status(rid)
    assert st2["steps"] == steps, st2
    sec = t1 - t0
    # per-kernel durations with HIP events on the pool's stream
    tv[0] = st2["dt"]
    tv[7] = kernel_steps
    pool.begin(rid, tv)
    kt = pool.time_kernels(kernel_steps)
    pos = pool.download(rid, pos1=False, tangent=False, forces=False)["pos"]
    pool.close()
    return dict(sec=sec, bead_steps_per_s=n * steps / sec, ms_per_step=1e3 * sec / steps,
                kernels_ms=kt, replays=st2["replays"], dt=st2["dt"], steps_per_graph=batch,
                rebuild_every=rebuild, finite=bool(np.all(np.isfinite(pos))))


def host_cores():
    """Threads for the CPU baseline: every core this process may run on
    (sched_getaffinity), or fewer where OMP_NUM_THREADS states the host's
    share for this job; with the CPU model and the machine's core count."""
    granted = len(os.sched_getaffinity(0))
    threads = granted
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and 0 < int(omp) < granted:
        threads = int(omp)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(threads=threads, affinity=granted, nproc=os.cpu_count(), cpu_model=model,
                omp_num_threads=omp or None)


def cpu_baseline(w, budget_s):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_lib as O
    hc = host_cores()
    threads = hc["threads"]
    sysd = dict(dim=2, n=w.n, n_nearby=w.phys.n_nearby, has_springs=True, has_wca=True,
                has_propel=True, L=w.phys.L, D=w.phys.D, Dr=w.phys.Dr, radius=w.phys.radius,
                wca=np.asarray(w.phys.wca), spring=np.asarray(w.phys.spring), v0=w.phys.v0)
    tv = w.time_vals.copy()
    pos, tang = w.pos, w.tangent
    done, sec = 0, 0.0
    # chunks of 40 steps (one Verlet rebuild each, the reference's cadence)
    # until about `budget_s` seconds of CPU work
    while sec < budget_s and done < 100000:
        tv[7] = 40
        t0 = time.perf_counter()
        pos, _, tang, _, st = O.srk2_run(sysd, pos, tang, w.chain_id, w.rest_lengths,
                                         w.v_tan_angle, tv, xi=None, seed=1, step0=done,
                                         pair_mode=1, n_threads=threads)
        sec += time.perf_counter() - t0
        tv[0] = st[0]
        done += 40
    return dict(value=w.n * done / sec, unit="bead-steps/s", cores=threads, kind="port",
                host=hc,
                sample=f"{done} SRK2 steps of the full {w.n}-bead system ({sec:.1f} s), C/OpenMP "
                       f"FP64 restatement of the reference algorithm (cell grid + Verlet list "
                       f"rebuilt every 40 steps, Newton-3 inside each thread chunk)")


def ensemble_part(args):
    """The 256-replica ensemble (chain counts log-U[50, 2000], lengths U{5..40},
    chirality U[-0.5, 0.5], v0 U[0.1, 2], phi U[0.1, 0.5], WCA, mixed
    precision, dt 1e-3) in one pool with its tiling and re-sort period tuned
    on the pool, `ens_steps` steps each after a warm-up; against a random
    sample of the replicas run one at a time in their own pools with the same
    settings, each warmed up (graph capture) before its clock starts."""
    from gcr_amd import engine as E
    from gcr_amd import ensemble as ENS
    from gcr_amd import workloads as W
    t0 = time.perf_counter()
    ws = W.uneven_replicas(seed=0, n_replicas=args.replicas, steps=args.ens_steps)
    t_init = time.perf_counter() - t0
    # the skin (each candidate a pool of its own, tuned by tune_pool), then
    # the pool of the chosen skin tuned again and timed
    t0 = time.perf_counter()
    frac, _, skin_trials = ENS.tune_skin(ws, E.GCR_MIXED, batch=args.batch)
    ws = ENS.with_skin(ws, frac)
    pool = ENS.pool_for(ws, E.GCR_MIXED)
    try:
        rids = ENS.add_all(pool, ws)
        tuned = ENS.tune_pool(pool, ws, rids, batch=args.batch)
        tuned["skin_frac"] = frac
        tuned["skin_trials"] = skin_trials
        t_tune = time.perf_counter() - t0
        kb = tuned["steps_per_batch"]
        for w, r in zip(ws, rids):  # warm-up: the graphs of the tuned schedule
            tv = w.time_vals.copy()
            tv[7] = kb
            pool.begin(r, tv)
        pool.run()
        for w, r in zip(ws, rids):
            tv = w.time_vals.copy()
            tv[0] = pool.status(r)["dt"]
            tv[7] = args.ens_steps
            pool.begin(r, tv)
        t0 = time.perf_counter()
        pool.run()
        sec = time.perf_counter() - t0
        sts = [pool.status(r) for r in rids]
        assert all(st["steps"] == args.ens_steps for st in sts)
        replays = int(sum(st["replays"] for st in sts))
    finally:
        pool.close()
    bs = sum(w.n for w in ws) * args.ens_steps
    rng = np.random.default_rng(1)
    sample = [ws[i] for i in rng.choice(len(ws), size=min(args.seq_sample, len(ws)), replace=False)]
    # one at a time: at the pool's tuned settings and at the defaults (skin 0.3
    # cutoff, re-sort every 16 steps, automatic tiles), whichever is faster
    ssec, sbs = ENS.run_sequential(sample, args.seq_steps, steps_per_batch=tuned["steps_per_batch"],
                                   rebuild_every=tuned["rebuild_every"], tile=tuned["tile"],
                                   threads=tuned["threads"])
    dsec, dbs = ENS.run_sequential(ENS.with_skin(sample, 0.3), args.seq_steps, steps_per_batch=args.batch,
                                   rebuild_every=16)
    seq_how = "the pool's tuned settings"
    if dbs / dsec > sbs / ssec:
        seq_tuned = sbs / ssec
        ssec, sbs = dsec, dbs
        seq_how = "the default settings (skin 0.3 cutoff, re-sort every 16 steps, automatic tiles)"
    else:
        seq_tuned = sbs / ssec
    value = bs / sec
    # what one persistent launch per batch would pay instead of the graph's
    # back-to-back launches: a grid-wide barrier at each boundary between the
    # step's kernels (stage 0 | stage 1 | max force and dt | update: 3 per
    # step), at the residency of the stage kernels (four 256-thread tiles
    # per CU)
    from gcr_amd import _lib
    bar_us, bar_blocks = _lib.grid_barrier_us(4, 200)
    return {"metric": "ensemble aggregate bead-steps/sec (256 uneven replicas, 1 GPU)",
            "value": value, "unit": "bead-steps/s", "steps": args.ens_steps,
            "ms_per_step": 1e3 * sec / args.ens_steps, "seconds": sec,
            "dtype": "f32 forces + f64 state",
            "config": {"workload": "256 replicas: chain counts log-U[50,2000], lengths U{5..40}, "
                                   "chirality U[-0.5,0.5], v0 U[0.1,2], phi U[0.1,0.5], WCA, "
                                   "dt=1e-3, mixed precision", **ENS.summary(ws),
                       "tuned": tuned, "tuning_s": t_tune},
            "sequential": {"value": sbs / ssec, "replicas_timed": len(sample),
                           "steps_each": args.seq_steps,
                           "at_tuned_settings": seq_tuned, "at_default_settings": dbs / dsec,
                           "how": "each replica alone in its own pool, one graph batch of warm-up "
                                  "before its clock starts, at " + seq_how + " (the faster of "
                                  "the pool's tuned settings and the defaults)"},
            "batched_over_sequential": value / (sbs / ssec),
            "roofline": {"bound": "hbm", "achieved": 80 * value / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": 80 * value / 1e9 / HBM_PEAK_GBS,
                         "note": "80 algorithmic bytes per 2D chain bead-step"},
            "persistent_launch": {"grid_barrier_us": bar_us, "workgroups": bar_blocks,
                                  "barriers_per_step": 3,
                                  "barrier_us_per_step": 3 * bar_us,
                                  "graph_ms_per_step": 1e3 * sec / args.ens_steps,
                                  "note": "a persistent kernel would replace the graph's "
                                          "back-to-back launches (no gap between them in the "
                                          "kernel trace) by these barriers"},
            "replays": replays, "host_init_s": t_init}


def main_ensemble(args):
    """The ensemble part alone, as one JSON line."""
    out = ensemble_part(args)
    out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "data": "synthetic"})
    print(json.dumps(out), flush=True)


def serving_part(args):
    """A Poisson stream (`rate` requests/s, 20 by default) of `requests`
    simulations through the continuous-batching server under an HBM budget:
    bead counts log-U[1k, 200k], step counts log-U[1e3, 1e5] times
    steps_scale, ABPs (D_r = 0.1) and chains of 5-40 beads, 2D and 3D, mixed
    precision. At 20 requests/s the stream offers about ten times the work
    one GPU does, so the latencies are those of a growing queue; `offered`
    against `value` says how loaded the server was."""
    from gcr_amd import serving as SV
    t0 = time.perf_counter()
    reqs = SV.make_stream(seed=0, n_requests=args.requests, rate=args.rate,
                          scale_steps=args.steps_scale)
    t_init = time.perf_counter() - t0
    srv = SV.Server(hbm_budget_bytes=int(args.budget_gb * (1 << 30)),
                    steps_per_batch=max(args.batch, args.serve_rebuild), rebuild_every=args.serve_rebuild,
                    skin_frac=args.serve_skin if args.serve_skin > 0 else None)
    try:
        done = srv.serve(reqs)
        graphs = {f"{d}d": lane.pool.graph_stats() for d, lane in srv.lanes.items()}
        lanes = {f"{d}d": {"batches": lane.batches, "busy_s": lane.busy_s}
                 for d, lane in srv.lanes.items()}
    finally:
        srv.close()
    rep = SV.report(done)
    work = sum(r.n * r.steps for r in reqs)
    offered = work / max(r.arrival for r in reqs)
    return {"metric": "serving aggregate bead-steps/sec + p50/p99 completion latency (1 GPU)",
            "value": rep["bead_steps_per_s"], "unit": "bead-steps/s",
            "dtype": "f32 forces + f64 state",
            "config": {"workload": f"Poisson stream, {args.rate:g} requests/s; beads log-U[1k,200k]; "
                                   "steps log-U[1e3,1e5] x steps_scale; ABPs (D_r=0.1) and chains "
                                   "5-40, 2D and 3D; mixed precision",
                       "requests": len(reqs), "rate_per_s": args.rate,
                       "steps_scale": args.steps_scale, "rebuild_every": args.serve_rebuild,
                       "skin_frac": args.serve_skin if args.serve_skin > 0 else 0.3,
                       "hbm_budget_gb": args.budget_gb, "total_bead_steps": work},
            "p50_latency_s": rep["p50_latency_s"], "p99_latency_s": rep["p99_latency_s"],
            "mean_queue_s": rep["mean_queue_s"], "makespan_s": rep["makespan_s"],
            "failed": rep["failed"], "offered_bead_steps_per_s": offered,
            "graphs": graphs, "lanes": lanes, "host_init_s": t_init}


def main_serve(args):
    """The serving part alone, as one JSON line."""
    out = serving_part(args)
    out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "data": "synthetic"})
    print(json.dumps(out), flush=True)


def cpu_option(w, budget_s):
    """The C/OpenMP FP64 oracle's SRK loop (Verlet list rebuilt every 40
    steps, Newton-3 inside each thread chunk) on the single-population
    workload w with all of its physics options (SRK order, FENE, walls, rigid
    rod, active sliding over the Verlet list with w.chunks per chain), in
    chunks of 40 steps until about budget_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_lib as O
    threads = host_cores()["threads"]
    p = w.phys
    sysd = dict(dim=p.dim, n=w.n, n_nearby=p.n_nearby, has_springs=p.has_springs,
                has_wca=p.has_wca, has_propel=p.has_propel, L=p.L, D=p.D, Dr=p.Dr,
                radius=p.radius, wca=np.asarray(p.wca), spring=np.asarray(p.spring), v0=p.v0,
                rod=p.rod, order=p.srk_order, walls=p.walls, wall_zone=p.wall_zone,
                wall_cutoff=p.wall_cutoff)
    if p.sliding_maxdist > 0:  # the chain chunks per bead, where_active
        sysd.update(sliding_maxdist=p.sliding_maxdist, sliding_v0=p.sliding_v0,
                    chunk=np.ascontiguousarray(w.chunks[w.chain_id], dtype=np.int32),
                    where_active=np.zeros(w.n, dtype=np.int32))
    tv = w.time_vals.copy()
    pos, tang = w.pos, w.tangent
    done, sec = 0, 0.0
    while sec < budget_s and done < 100000:
        tv[7] = 40
        t0 = time.perf_counter()
        pos, _, tang, _, st = O.srk2_run(sysd, pos, tang, w.chain_id, w.rest_lengths,
                                         w.v_tan_angle, tv, xi=None, seed=1, step0=done,
                                         pair_mode=1, n_threads=threads)
        sec += time.perf_counter() - t0
        tv[0] = st[0]
        done += 40
    return dict(bead_steps_per_s=w.n * done / sec, steps=done, seconds=sec, threads=threads)


def cpu_growth(final, pos, tangent, chain_id, chain_len, rest_length, time_vals, t_init_grow,
               stepskip, budget_s):
    """The growth schedule (control.py:153-219) driven through the C/OpenMP
    FP64 oracle chunk by chunk (Verlet lists, rebuilt every chunk) until the
    schedule ends or about budget_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_lib as O
    from gcr_amd import grow as G
    threads = host_cores()["threads"]
    tv = np.array(time_vals, dtype=np.float64).copy()
    tv[3], tv[4] = 1.0, -1.0
    tv[5] = tv[6] = tv[7] = stepskip
    vt = np.zeros(final.n_chains)
    sched = G.growth_schedule(final, chain_len, rest_length, t_init_grow)
    done, sec, step = 0, 0.0, next(sched, None)
    while step is not None and sec < budget_s:
        ph, rest = step
        sysd = dict(dim=ph.dim, n=ph.n_beads, n_nearby=ph.n_nearby, has_springs=ph.has_springs,
                    has_wca=ph.has_wca, has_propel=ph.has_propel, L=ph.L, D=ph.D, Dr=ph.Dr,
                    radius=ph.radius, wca=np.asarray(ph.wca), spring=np.asarray(ph.spring),
                    v0=ph.v0, rod=ph.rod)
        t0 = time.perf_counter()
        pos, _, tangent, _, st = O.srk2_run(sysd, pos, tangent, chain_id, rest, vt, tv, seed=1,
                                            step0=done, pair_mode=1, n_threads=threads)
        sec += time.perf_counter() - t0
        tv[0] = st[0]
        done += int(st[2])
        try:
            step = sched.send(st[1])
        except StopIteration:
            step = None
    return dict(bead_steps_per_s=final.n_beads * done / sec, steps=done, seconds=sec,
                threads=threads)


def cpu_groups(reps, laws, fixed, budget_s):
    """The C/OpenMP FP64 oracle's multi-population SRK2 loop on the
    populations `reps` (Workloads) with the laws {(P, Q): WCA vector or
    "motor"} and the fixed indices: each population's own pairs over its
    Verlet list, the other populations and the motors through cell grids,
    in chunks of 40 steps until about budget_s seconds."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_lib as O
    threads = host_cores()["threads"]
    pops = []
    for k, w in enumerate(reps):
        p = w.phys
        sysd = dict(dim=p.dim, n=w.n, n_nearby=p.n_nearby, has_springs=p.has_springs,
                    has_wca=p.has_wca, has_propel=p.has_propel, L=p.L, D=p.D, Dr=p.Dr,
                    radius=p.radius, wca=np.asarray(p.wca), spring=np.asarray(p.spring),
                    v0=p.v0, where_active=np.zeros(w.n, dtype=np.int32))
        d = dict(sysd=sysd, pos=w.pos, tangent=w.tangent, chain_id=w.chain_id,
                 rest_lengths=w.rest_lengths, vtan=w.v_tan_angle, seed=p.seed,
                 fixed=k in fixed)
        if p.motor:
            d["motor"] = dict(max_dist=p.motor_max_dist, imparts=p.motor_imparts, k_on=p.k_on,
                              k_off=p.k_off, reach=p.motor_reach or p.motor_max_dist,
                              still=p.motor_still)
        pops.append(d)
    moving = [k for k in range(len(reps)) if k not in fixed]
    tv = reps[moving[0]].time_vals.copy()
    done, sec = 0, 0.0
    while sec < budget_s and done < 100000:
        tv[7] = 40
        t0 = time.perf_counter()
        out, st = O.multi_srk2_run(pops, laws, tv, step0=done, pair_mode=1, n_threads=threads)
        sec += time.perf_counter() - t0
        for k in moving:
            pops[k]["pos"], pops[k]["tangent"] = out[k][0], out[k][2]
        tv[0] = st[0]
        done += 40
    moved = sum(reps[k].n for k in moving)
    return dict(bead_steps_per_s=moved * done / sec, steps=done, seconds=sec, threads=threads)


def time_pool(reps, steps, warmup, precision, dim=2, group=None, chunks=None, batch=64,
              rebuild=32, tile=8, threads=256):
    """Bead-steps/s of the replicas `reps` (Workloads) advanced together in one
    pool (a group when `group` = (laws, fixed indices) is given), timed over
    `steps` steps after `warmup`; the replays the run took."""
    from gcr_amd import engine as E
    nb = sum(w.n for w in reps)
    cells = sum(E.replica_cells(w.phys) for w in reps)
    pool = E.Pool(dim, precision, bead_capacity=nb, cell_capacity=cells + len(reps),
                  chain_capacity=sum(w.phys.n_chains for w in reps),
                  replica_capacity=len(reps))
    if dim == 3:  # the 2D tile settings do not carry over: 3D tiles are picked per replica
        tile, threads = 0, 0
    try:
        pool.set_schedule(batch, rebuild)
        pool.set_tile(tile)
        pool.set_tile_threads(threads)
        rids = [pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
                for w in reps]
        fixed = set()
        if group is not None:
            laws, fixed = group
            pool.group(rids, laws, sorted(fixed))
        for k, c in (chunks or {}).items():
            pool.set_chunks(rids[k], c)
        moving = [k for k in range(len(reps)) if k not in fixed]
        tv = reps[moving[0]].time_vals.copy()
        tv[7] = warmup
        for k in moving:
            pool.begin(rids[k], tv)
        pool.run()
        tv[0] = pool.status(rids[moving[0]])["dt"]
        tv[7] = steps
        for k in moving:
            pool.begin(rids[k], tv)
        t0 = time.perf_counter()
        pool.run()
        sec = time.perf_counter() - t0
        sts = [pool.status(rids[k]) for k in moving]
        assert all(s["steps"] == steps for s in sts), sts
        moved = sum(reps[k].n for k in moving)
        return dict(bead_steps_per_s=moved * steps / sec, ms_per_step=1e3 * sec / steps,
                    beads=moved, replays=int(sum(s["replays"] for s in sts)))
    finally:
        pool.close()


def main_features(args):
    """Throughput of each option of the step on the 400k-bead system (mixed
    precision, the tiling the autotuner picks for the plain step): SRK4,
    FENE springs, walls, the rigid rod, active sliding, 3D chains (100k
    beads), two interacting populations, chains with motors, and the growth
    phase on 40k beads."""
    import copy
    from gcr_amd import dropin
    from gcr_amd import engine as E
    from gcr_amd import grow as G
    from gcr_amd import workloads as W
    prec = E.GCR_MIXED
    base = W.chains_20kx20(seed=1, n_chains=args.chains)
    base.phys.skin = 0.3 * float(base.phys.wca[3])
    steps, warm = args.steps, args.warmup
    out = {}

    def variant(**over):
        w = copy.deepcopy(base)
        for k, v in over.items():
            setattr(w.phys, k, v)
        return w

    def fene_w():
        w = variant()
        w.phys.spring = w.phys.spring.copy()
        w.phys.spring[3], w.phys.spring[5] = 1.0, 0.3
        return w

    def sliding_w():
        w = variant(has_propel=False, sliding_maxdist=1.5, sliding_v0=1.0)
        w.chunks = dropin.chain_chunks(3, w.chain_id)  # three threads' chunks
        return w

    def chains3d_w():
        w = W.chains("chains3d", 3, 5000, 20, phi=0.15, cutoff_factor=W.LJ_25_SIGMA, kappa=5.0,
                     T=0.05, dtmax=5e-4, seed=4)
        w.phys.skin = 0.3 * float(w.phys.wca[3])
        return w

    # single-population options: (workload, algorithmic HBM bytes per
    # bead-step: x read, x1 written, x1 read, x read, x written, 8 B per
    # coordinate; SRK4 moves positions through twice as many passes)
    single = {
        "srk2 (reference step)": (lambda: base, 80),
        "srk4": (lambda: variant(srk_order=4), 160),
        "fene": (fene_w, 80),
        "walls (x and y)": (lambda: W.walled(base, 3, 1.25), 80),
        "rigid rod 0.6": (lambda: variant(rod=0.6), 80),
        "active sliding": (sliding_w, 80),
        # 3D: chains of 20 at volume fraction 0.15 with the same law (120 B
        # per bead-step: three coordinates)
        "3D chains, 5000 x 20": (chains3d_w, 120),
    }

    def two_pops_set():
        # the same beads as two populations of 10k chains acting on each
        # other with their own law (the union of the reference step)
        half = args.chains // 2
        parts = []
        for lo, hi in ((0, half), (half, args.chains)):
            w = copy.deepcopy(base)
            sl = slice(lo * 20, hi * 20)
            w.pos, w.tangent = base.pos[sl].copy(), base.tangent[sl].copy()
            w.chain_id = np.repeat(np.arange(hi - lo, dtype=np.int64), 20)
            w.rest_lengths = base.rest_lengths[lo:hi].copy()
            w.v_tan_angle = base.v_tan_angle[lo:hi].copy()
            w.v0_per_chain = base.v0_per_chain[lo:hi].copy()
            w.phys.n_chains, w.phys.n_beads = hi - lo, (hi - lo) * 20
            w.phys.seed = 100 + lo
            parts.append(w)
        law = np.asarray(base.phys.wca)
        return parts, {(0, 1): law, (1, 0): law}, set()

    def motors_set():
        # chains whose v0 is the motors' push (config.py:95-98) among 100k fixed motors
        chains = variant(v0=0.0)
        rng = np.random.default_rng(5)
        nm = 100000
        mot = copy.deepcopy(base)
        mot.pos = rng.uniform(0, base.phys.L, (nm, 2))
        mot.tangent = np.tile([1.0, 0.0], (nm, 1))
        mot.chain_id = np.arange(nm, dtype=np.int64)
        mot.rest_lengths = np.ones(nm)
        mot.v_tan_angle, mot.v0_per_chain = np.zeros(nm), np.zeros(nm)
        mot.phys = E.Physics(dim=2, n_beads=nm, n_chains=nm, L=base.phys.L, D=0.0, Dr=0.0,
                             radius=1.0, wca=np.array([0.0, 0.0, 0.0, 1.0]),
                             spring=np.zeros(6), v0=0.0, n_nearby=-1, has_springs=False,
                             has_wca=False, has_propel=False, skin=0.3, motor=True,
                             motor_max_dist=0.25, motor_imparts=1.0, k_on=0.9, k_off=0.2,
                             seed=9)
        return [chains, mot], {(0, 1): "motor"}, {1}

    groups = {"two populations": two_pops_set, "chains + 100k fixed motors": motors_set}

    def growth_set():
        # the growth schedule (control.py:153-219) for 2000 chains of 20 from points
        g = W.chains_20kx20(seed=2, n_chains=2000)
        pos0, t0_, cid = G.random_initialization(2000, 20, 0.5, g.phys.L, 2, seed=3)
        return g, pos0, t0_, cid

    def growth():
        g, pos0, t0_, cid = growth_set()
        tg = time.perf_counter()
        _, _, info = G.grow(g.phys, pos0, t0_, cid, 0.5, 20, g.time_vals, t_init_grow=0.2,
                            growing_stepskip=10)
        sec = time.perf_counter() - tg
        return dict(bead_steps_per_s=g.n * info["steps"] / sec,
                    ms_per_step=1e3 * sec / info["steps"], beads=g.n, steps=info["steps"],
                    chunks=info["chunks"])

    def single_run(mk):
        w = mk()
        chunks = {0: w.chunks} if getattr(w, "chunks", None) is not None else None
        return time_pool([w], steps, warm, prec, dim=w.phys.dim, chunks=chunks)

    options = {name: (lambda mk=mk: single_run(mk)) for name, (mk, _) in single.items()}
    def group_run(mk):
        reps, laws, fixed = mk()
        return time_pool(reps, steps, warm, prec, group=(laws, fixed))

    for name, mk in groups.items():
        options[name] = (lambda mk=mk: group_run(mk))
    options["growth, 2000 chains x 20 (FP64)"] = growth
    t0 = time.perf_counter()
    for name, run in options.items():
        if args.only not in name:
            continue
        r = run()
        if name in single:
            # the whole step against HBM's 8 TB/s at the algorithmic bytes
            bps = single[name][1]
            gbs = bps * r["bead_steps_per_s"] / 1e9
            r["step_roofline"] = dict(bound="hbm", bytes_per_bead_step=bps,
                                      achieved=gbs, peak=8000.0, unit="GB/s", frac=gbs / 8000.0)
            if not args.no_cpu:
                c = cpu_option(single[name][0](), args.cpu_seconds / 2)
                r["cpu_baseline"] = dict(value=c["bead_steps_per_s"], unit="bead-steps/s",
                                         cores=c["threads"], kind="port",
                                         sample=f"{c['steps']} steps of the same system "
                                                f"({c['seconds']:.1f} s)")
                r["gpu_over_cpu"] = r["bead_steps_per_s"] / c["bead_steps_per_s"]
        elif name in groups:
            gbs = 80 * r["bead_steps_per_s"] / 1e9
            r["step_roofline"] = dict(bound="hbm", bytes_per_bead_step=80, achieved=gbs,
                                      peak=8000.0, unit="GB/s", frac=gbs / 8000.0)
            if not args.no_cpu:
                reps, laws, fixed = groups[name]()
                c = cpu_groups(reps, laws, fixed, args.cpu_seconds / 2)
                r["cpu_baseline"] = dict(value=c["bead_steps_per_s"], unit="bead-steps/s",
                                         cores=c["threads"], kind="port",
                                         sample=f"{c['steps']} steps of the same populations "
                                                f"({c['seconds']:.1f} s)")
                r["gpu_over_cpu"] = r["bead_steps_per_s"] / c["bead_steps_per_s"]
        elif not args.no_cpu:
            g, pos0, t0_, cid = growth_set()
            c = cpu_growth(g.phys, pos0, t0_, cid, 20, 0.5, g.time_vals, 0.2, 10,
                           args.cpu_seconds / 2)
            r["cpu_baseline"] = dict(value=c["bead_steps_per_s"], unit="bead-steps/s",
                                     cores=c["threads"], kind="port",
                                     sample=f"the first {c['steps']} steps of the same schedule "
                                            f"({c['seconds']:.1f} s)")
            r["gpu_over_cpu"] = r["bead_steps_per_s"] / c["bead_steps_per_s"]
        out[name] = r
    res = {"metric": "bead-steps/sec of each step option (1 GPU)", "unit": "bead-steps/s",
           "dtype": "f32 forces + f64 state", "data": "synthetic", "steps": steps,
           "warmup": warm, "config": {"workload": "20k chains x 20 beads (400k) and its variants",
                                      "skin": base.phys.skin, "rebuild_every": 32, "tile": 8,
                                      "threads": 256},
           "features": out, "wall_s": time.perf_counter() - t0}
    print(json.dumps(res), flush=True)


def main_observables(args):
    """The observables after the step on the 400k-bead system: polar and
    nematic order with the director (analysis.py:294-328 over the per-chain
    mean angles, output_methods.py:492-510) on the device, against the host
    formulas on host arrays (the reference's arrays live on the host); and
    one .dat snapshot (output_methods.py:420-485), download included."""
    import tempfile
    from gcr_amd import engine as E
    from gcr_amd import observables as OB
    from gcr_amd import snapshot as SN
    from gcr_amd import workloads as W
    w = W.chains_20kx20(seed=1, n_chains=args.chains)
    w.phys.skin = 0.3 * float(w.phys.wca[3])
    pool = E.Pool(2, E.GCR_MIXED, bead_capacity=w.n, cell_capacity=E.replica_cells(w.phys) + 1,
                  chain_capacity=w.phys.n_chains, replica_capacity=1)
    try:
        rid = pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        tv = w.time_vals.copy()
        tv[7] = 64
        pool.begin(rid, tv)
        pool.run()
        calls = 200
        for _ in range(10):
            dev = pool.order_parameters(rid)
        t0 = time.perf_counter()
        for _ in range(calls):
            dev = pool.order_parameters(rid)
        dev_s = (time.perf_counter() - t0) / calls
        t0 = time.perf_counter()
        out = pool.download(rid, pos1=False, forces=False)
        dl_s = time.perf_counter() - t0
    finally:
        pool.close()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        phi = OB.chain_mean_angle(out["tangent"], w.chain_id, w.phys.n_chains)
        host = (OB.polar_order(phi),) + OB.nematic_order(phi)
    host_s = (time.perf_counter() - t0) / reps
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        fn = SN.save_state(os.path.join(d, "run"), "chains", 1, 64, 0.1, out["pos"],
                           out["tangent"], w.chain_id)
        snap_s = time.perf_counter() - t0
        size = os.path.getsize(fn)
    # the order parameters read one tangent per bead (16 B) and write a
    # few numbers: algorithmic bytes 16 per bead
    gbs = 16 * w.n / dev_s / 1e9
    res = {"metric": "observables of the 400k-bead system (1 GPU)", "unit": "s per call",
           "higher_is_better": False, "data": "synthetic",
           "config": {"workload": "20k chains x 20 beads (400k) after 64 mixed steps"},
           "order_parameters": {
               "device_s": dev_s, "host_numpy_s": host_s, "device_over_host": host_s / dev_s,
               "polar": dev["polar"], "nematic": dev["nematic"],
               "host_polar": host[0], "host_nematic": host[1],
               "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                            "frac": gbs / 8000.0,
                            "note": "whole call, launch and host synchronisation included"},
               "cpu_baseline": {"value": host_s, "unit": "s per call", "cores": 1,
                                "kind": "port", "sample": f"{reps} calls of numpy on host "
                                                         "arrays"}},
           "snapshot": {"download_s": dl_s, "write_s": snap_s, "bytes": size,
                        "write_MB_per_s": size / snap_s / 1e6}}
    print(json.dumps(res), flush=True)


def main_copy():
    """dst = 1 * src over 1 GiB of float64, three times: one elementwise
    kernel that reads 1 GiB and writes 1 GiB per launch. Counter readings of
    this kernel convert FETCH_SIZE / WRITE_SIZE into bytes."""
    import torch
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "copy", "bytes_read_per_launch": 8 * n,
                      "bytes_written_per_launch": 8 * n, "launches": 3}), flush=True)


def main_request(args):
    """Requests of the serving stream (seed 0, --requests, --rate), each alone
    in a pool at the server's settings (skin --serve-skin of its cutoff,
    re-sort every --serve-rebuild steps, batches of --batch): 128 warm-up
    steps, --steps timed steps, then the per-kernel durations over
    --kernel-steps steps. One JSON line per request."""
    import copy
    from gcr_amd import engine as E
    from gcr_amd import serving as SV
    reqs = SV.make_stream(seed=0, n_requests=args.requests, rate=args.rate)
    for i in (int(x) for x in args.ids.split(",")):
        r = reqs[i]
        w = copy.copy(r.work)
        w.phys = copy.copy(r.work.phys)
        if args.serve_skin > 0:
            w.phys.skin = args.serve_skin * float(w.phys.wca[3])
        pool = E.Pool(w.phys.dim, E.GCR_MIXED, w.n, E.replica_cells(w.phys) + 1,
                      w.phys.n_chains, 1)
        try:
            batch = max(args.batch, args.serve_rebuild)
            pool.set_schedule(batch, args.serve_rebuild)
            rid = pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                           w.v0_per_chain)
            tv = w.time_vals.copy()
            tv[7] = max(128, batch)
            pool.begin(rid, tv)
            pool.run()
            tv[0] = pool.status(rid)["dt"]
            tv[7] = args.steps
            pool.begin(rid, tv)
            t0 = time.perf_counter()
            pool.run()
            sec = time.perf_counter() - t0
            st = pool.status(rid)
            assert st["steps"] == args.steps, st
            tv[0] = st["dt"]
            tv[7] = args.kernel_steps
            pool.begin(rid, tv)
            kt = pool.time_kernels(args.kernel_steps)
        finally:
            pool.close()
        print(json.dumps({"mode": "request", "request": i, "dim": r.dim, "beads": r.n,
                          "chain_length": r.n // max(1, w.phys.n_chains), "steps": args.steps,
                          "bead_steps_per_s": r.n * args.steps / sec,
                          "ms_per_step": 1e3 * sec / args.steps,
                          "kernels_ms": {k: float(v) for k, v in kt.items()},
                          "request_steps": r.steps, "alone_s": r.steps * sec / args.steps}),
              flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.mode == "copy":
        return main_copy()
    if args.mode == "ensemble":
        return main_ensemble(args)
    if args.mode == "serve":
        return main_serve(args)
    if args.mode == "request":
        return main_request(args)
    if args.mode == "features":
        return main_features(args)
    if args.mode == "observables":
        return main_observables(args)
    from gcr_amd import engine as E
    from gcr_amd import workloads as W
    from gcr_amd._lib import device_info
    w = W.chains_20kx20(seed=1, n_chains=args.chains)
    dev = device_info()
    res, tuned = {}, {}
    for name, prec in (("fp64", E.GCR_FP64), ("mixed", E.GCR_MIXED)):
        if args.no_autotune:
            cfg = dict(skin=args.skin if args.skin > 0 else 0.3 * float(w.phys.wca[3]),
                       rebuild_every=args.rebuild, tile=parse_tile(args.tile), threads=args.threads)
        else:
            from gcr_amd import autotune as AT
            t0 = time.perf_counter()
            t = AT.tune(w, prec)
            cfg = {k: t[k] for k in ("skin", "rebuild_every", "tile", "threads")}
            cfg["tuning_s"] = time.perf_counter() - t0
            cfg["tiling_top"] = [[r["tile"], r["threads"], round(1e3 * r["stages_ms"], 1)]
                                 for r in t["tiling_trials"][:4]]
            cfg["schedule_top"] = [[round(r["skin"], 3), r["rebuild_every"], round(1e3 * r["ms_per_step"], 1),
                                    r["replays"]]
                                   for r in sorted(t["schedule_trials"], key=lambda r: r["ms_per_step"])[:5]]
        tuned[name] = cfg
        w.phys.skin = cfg["skin"]
        res[name] = run_gpu(w, prec, args.steps, args.warmup, args.batch, cfg["rebuild_every"],
                            args.kernel_steps, cfg["tile"], cfg["threads"])
    m = res["mixed"]
    kt = m["kernels_ms"]
    dom = max(("stage0", "stage1", "update"), key=lambda k: kt[k])
    achieved = ALG_BYTES[dom] * w.n / (kt[dom] * 1e-3) / 1e9
    out = {
        "metric": "bead-steps/sec (400k beads, 1 GPU), mixed vs FP64; ensemble agg. + p50/p99 "
                  "latency",
        "value": m["bead_steps_per_s"],
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces + f64 state",
        "data": "synthetic (random non-overlapping chains from a seed)",
        "config": {"workload": "2D periodic, 20k chains x 20 beads (400k), phi=0.4, LJ eps=1 "
                               "cutoff 2.5 sigma, kappa=5, chirality U[-0.5,0.5]/chain, v0=1, "
                               "T=0.05, dtmax=5e-4, SRK2",
                   "beads": w.n, "box": w.phys.L, "steps_per_graph": m["steps_per_graph"],
                   "rebuild_every": m["rebuild_every"],
                   "cell_list": tuned},
        "fp64": {"value": res["fp64"]["bead_steps_per_s"], "ms_per_step": res["fp64"]["ms_per_step"],
                 "kernels_ms": res["fp64"]["kernels_ms"]},
        "mixed_over_fp64": m["bead_steps_per_s"] / res["fp64"]["bead_steps_per_s"],
        "kernels_ms": kt,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES[dom] * w.n,
                     "traffic_gbs": TRAFFIC_BYTES[dom] * w.n / (kt[dom] * 1e-3) / 1e9,
                     "step_achieved": 80 * w.n / (m["ms_per_step"] * 1e-3) / 1e9},
        "device": dev[0],
        "replays": {k: v["replays"] for k, v in res.items()},
    }
    if not args.no_ensemble:
        out["ensemble"] = ensemble_part(args)
    if not args.no_serving:
        out["serving"] = serving_part(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
