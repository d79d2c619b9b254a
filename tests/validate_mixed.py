"""Long-run validation of the mixed form (FP32 pair forces, FP64 positions,
wrapping and RK2 combination) against the FP64 path and the C oracle. A
measurement script run on demand (pytest does not collect it):

    python tests/validate_mixed.py --part trajectory|statistics|rotation [--out FILE.json]

trajectory  400k beads (20k chains x 20, phi 0.4, LJ cut at 2.5 sigma, kappa 5,
            chirality U[-0.5, 0.5], T 0.05, dt 5e-4) from one initial state
            with in-kernel Philox noise, which the oracle reproduces: the
            position error of the mixed and of the FP64 GPU path against the
            oracle after 1, 10, 100, 1000, 2000, ..., 10^4 steps (all three run
            in the same segments), beside the distance between two runs with
            different noise seeds, the scale at which trajectories are
            decorrelated.
statistics  the same system for 10^4 steps, `--seeds` noise seeds per
            precision: chain-centre MSD at the end, polar and nematic order
            over the second half (analysis.py:294-328), mean and standard
            error over the seeds, and z = (mixed - FP64) / (combined error).
rotation    1000 chains x 20, phi 0.3, WCA, kappa 10, chirality 0.5, T 0.01,
            dt 1e-3, from an aligned nematic start (rods along +-x), 10^5
            steps: the collective rotation rate of the director, the paper
            notebooks' estimator (from the first sample with nematic order
            > 0.7, director differences over 4 samples, folded mod pi, as
            the nematic director is headless) on the series that
            global_op-chiral.py:24-30 writes, one rate per seed
            (`--seeds` per precision), and one seed of the same system on
            the oracle, observables computed on the host from its arrays.

The GPU runs keep their observables on the device (gcr_replica_sample);
the oracle's are computed by gcr_amd.observables on its positions.
"""
import argparse
import copy
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "gliding-collective-rotation_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle_lib as O  # noqa: E402
from gcr_amd import engine as E  # noqa: E402
from gcr_amd import observables as OB  # noqa: E402
from gcr_amd import workloads as W  # noqa: E402

THREADS = len(os.sched_getaffinity(0))
T0 = time.perf_counter()


def log(msg):
    print(f"[{time.perf_counter() - T0:7.1f}s] {msg}", flush=True)


def sysd_of(phys):
    return dict(dim=phys.dim, n=phys.n_beads, n_nearby=phys.n_nearby,
                has_springs=phys.has_springs, has_wca=phys.has_wca, has_propel=phys.has_propel,
                L=phys.L, D=phys.D, Dr=phys.Dr, radius=phys.radius, wca=np.asarray(phys.wca),
                spring=np.asarray(phys.spring), v0=phys.v0)


class GpuRun:
    """One replica advanced in segments; observables sampled on the device."""

    def __init__(self, w, precision, seed, tile=0, threads=0, skin=None, rebuild=8):
        phys = copy.copy(w.phys)
        phys.seed = seed
        if skin is not None:
            phys.skin = skin
        self.pool = E.Pool(2, precision, bead_capacity=w.n,
                           cell_capacity=E.replica_cells(phys) + 1,
                           chain_capacity=phys.n_chains, replica_capacity=1)
        self.pool.set_schedule(64, rebuild)
        self.pool.set_tile(tile)
        self.pool.set_tile_threads(threads)
        self.rid = self.pool.add(phys, w.pos, w.tangent, w.chain_id, w.rest_lengths,
                                 w.v_tan_angle)
        self.tv = w.time_vals.copy()
        self.pool.sample(self.rid, reset=True)

    def advance(self, steps, sample=True):
        self.tv[7] = steps
        self.pool.begin(self.rid, self.tv)
        self.pool.run()
        st = self.pool.status(self.rid)
        if st["failed"] or st["steps"] != steps:
            raise RuntimeError(f"run failed: {st}")
        self.tv[0] = st["dt"]
        if sample:
            self.pool.sample(self.rid)
        return st

    def pos(self):
        return self.pool.download(self.rid, pos1=False, tangent=False, forces=False)["pos"]

    def series(self):
        return self.pool.series(self.rid)

    def close(self):
        self.pool.close()


class OracleRun:
    """The C oracle (Verlet-list path) advanced in the same segments, with
    the same counter-based noise (keyed by seed, bead and global step)."""

    def __init__(self, w, seed):
        self.w, self.seed = w, seed
        self.sysd = sysd_of(w.phys)
        self.p, self.t = w.pos.copy(), w.tangent.copy()
        self.tv = w.time_vals.copy()
        self.done = 0
        self.unw = OB.Unwrapper(self.p, w.phys.L)
        self.rows = []
        self.elapsed = 0.0
        self.sample()

    def advance(self, steps, sample=True):
        w = self.w
        self.tv[7] = steps
        p, _, t, _, st = O.srk2_run(self.sysd, self.p, self.t, w.chain_id, w.rest_lengths,
                                    w.v_tan_angle, self.tv, xi=None, seed=self.seed,
                                    step0=self.done, pair_mode=1, n_threads=THREADS)
        self.p, self.t = p, t
        self.tv[0] = st[0]
        self.elapsed += st[1]
        self.done += steps
        if sample:
            self.unw.add(p)
            self.sample()

    def sample(self):
        phi = OB.chain_mean_angle(self.t, self.w.chain_id)
        S, d = OB.nematic_order(phi)
        self.rows.append((self.elapsed, OB.polar_order(phi), S, d,
                          self.unw.chain_com_msd(self.w.chain_id)))

    def series(self):
        a = np.array(self.rows)
        return dict(t=a[:, 0], polar=a[:, 1], nematic=a[:, 2], director=a[:, 3], msd=a[:, 4])


def distance(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    r = np.sqrt(np.sum(d * d, axis=1))
    return dict(max=float(r.max()), rms=float(np.sqrt(np.mean(r * r))))


def bench_settings(w):
    rc = float(w.phys.wca[3])
    return dict(tile=8, threads=256, skin=0.3 * rc, rebuild=32)


def part_trajectory(args):
    w = W.chains_20kx20(seed=1)
    L = w.phys.L
    marks = [1, 10, 100] + list(range(1000, args.steps + 1, 1000))
    kw = bench_settings(w)
    mx = GpuRun(w, E.GCR_MIXED, w.phys.seed, **kw)
    f64 = GpuRun(w, E.GCR_FP64, w.phys.seed, **kw)
    other = GpuRun(w, E.GCR_FP64, w.phys.seed + 7919, **kw)
    orc = OracleRun(w, w.phys.seed)
    log(f"trajectory: {w.n} beads, L = {L:.3f}, oracle on {THREADS} threads")
    rows, done = [], 0
    for m in marks:
        n = m - done
        for r in (mx, f64, other):
            r.advance(n, sample=False)
        orc.advance(n, sample=False)
        done = m
        po, pm, pf, pb = orc.p, mx.pos(), f64.pos(), other.pos()
        row = dict(steps=m, mixed_vs_oracle=distance(pm, po, L),
                   fp64_vs_oracle=distance(pf, po, L), mixed_vs_fp64=distance(pm, pf, L),
                   other_noise_vs_fp64=distance(pb, pf, L))
        rows.append(row)
        log(f"step {m}: mixed-oracle max {row['mixed_vs_oracle']['max']:.3e} "
            f"rms {row['mixed_vs_oracle']['rms']:.3e}; fp64-oracle max "
            f"{row['fp64_vs_oracle']['max']:.3e} rms {row['fp64_vs_oracle']['rms']:.3e}; "
            f"other noise rms {row['other_noise_vs_fp64']['rms']:.3e}")
    for r in (mx, f64, other):
        r.close()
    p0, p1, rc = w.phys.wca[0], w.phys.wca[1], float(w.phys.wca[3])
    f_rc = abs(rc ** -4 * (p0 - p1 * rc ** -3)) * rc
    return dict(workload="400k beads: 20k chains x 20, phi 0.4, LJ 2.5 sigma, kappa 5, "
                "chirality U[-0.5,0.5], T 0.05, dt 5e-4", beads=w.n, box=L, rows=rows,
                cutoff_pair_step_bound=float(w.time_vals[1]) * f_rc / w.phys.radius)


def mean_se(x):
    x = np.asarray(x, dtype=np.float64)
    return float(x.mean()), float(x.std(ddof=1) / np.sqrt(len(x)))


def compare(per_seed):
    out = {}
    for k in per_seed[E.GCR_FP64][0]:
        a = [s[k] for s in per_seed[E.GCR_FP64]]
        b = [s[k] for s in per_seed[E.GCR_MIXED]]
        m64, e64 = mean_se(a)
        mmx, emx = mean_se(b)
        se = np.hypot(e64, emx)
        # the two precisions share each seed's noise, so the per-seed
        # difference is the sharper test
        dm, de = mean_se(np.subtract(b, a))
        out[k] = dict(fp64=m64, fp64_se=e64, mixed=mmx, mixed_se=emx,
                      z=float((mmx - m64) / se) if se > 0 else 0.0,
                      paired_diff=dm, paired_diff_se=de,
                      z_paired=float(dm / de) if de > 0 else 0.0,
                      fp64_seeds=a, mixed_seeds=b)
    return out


def part_statistics(args):
    w = W.chains_20kx20(seed=1)
    kw = bench_settings(w)
    every = 100
    per_seed = {E.GCR_FP64: [], E.GCR_MIXED: []}
    for i in range(args.seeds):
        for prec in (E.GCR_FP64, E.GCR_MIXED):
            r = GpuRun(w, prec, 1000 + i, **kw)
            for _ in range(args.steps // every):
                r.advance(every)
            s = r.series()
            r.close()
            h = len(s["t"]) // 2
            per_seed[prec].append(dict(msd_end=float(s["msd"][-1]),
                                       polar_2nd_half=float(s["polar"][h:].mean()),
                                       nematic_2nd_half=float(s["nematic"][h:].mean())))
            log(f"statistics seed {i} precision {prec}: {per_seed[prec][-1]} "
                f"t_end {s['t'][-1]:.3f}")
    return dict(workload="400k beads (as trajectory), noise seeds 1000..", steps=args.steps,
                sample_every=every, seeds=args.seeds, observables=compare(per_seed))


def rotation_workload(seed):
    w = W.chains("aligned_1000x20", 2, 1000, 20, phi=0.3, kappa=10.0, vtan=0.5, T=0.01,
                 dtmax=1e-3, seed=seed)
    return W.aligned_rods(w, seed=seed)


def rate_of(s):
    r = OB.rotation_rate(s["t"], s["director"], s["nematic"], threshold=0.7, lag=4,
                         period=np.pi)
    if r is None:
        return None
    return float(np.mean(r[1]))


def part_rotation(args):
    every = 1000
    per_seed = {E.GCR_FP64: [], E.GCR_MIXED: []}
    for i in range(args.seeds):
        w = rotation_workload(2000 + i)
        for prec in (E.GCR_FP64, E.GCR_MIXED):
            r = GpuRun(w, prec, w.phys.seed)
            for _ in range(args.rot_steps // every):
                r.advance(every)
            s = r.series()
            r.close()
            rate = rate_of(s)
            if rate is None:
                raise RuntimeError(f"seed {i}: never ordered")
            h = len(s["t"]) // 2
            per_seed[prec].append(dict(rotation_rate=rate,
                                       director_turn=float(OB.angle_difference(
                                           s["director"][-1], s["director"][0], np.pi)),
                                       nematic_2nd_half=float(s["nematic"][h:].mean()),
                                       msd_end=float(s["msd"][-1])))
            log(f"rotation seed {i} precision {prec}: {per_seed[prec][-1]} t_end {s['t'][-1]:.2f}")
    out = dict(workload="1000 chains x 20, phi 0.3, WCA, kappa 10, chirality 0.5, T 0.01, "
               "dt 1e-3, aligned nematic start", steps=args.rot_steps, sample_every=every,
               seeds=args.seeds, observables=compare(per_seed))
    if args.oracle_seeds:
        w = rotation_workload(2000)
        orc = OracleRun(w, w.phys.seed)
        for k in range(args.rot_steps // every):
            orc.advance(every)
            if k % 10 == 9:
                log(f"oracle rotation run: {orc.done} steps")
        s = orc.series()
        rate = rate_of(s)
        h = len(s["t"]) // 2
        o = dict(rotation_rate=rate, nematic_2nd_half=float(s["nematic"][h:].mean()),
                 msd_end=float(s["msd"][-1]))
        f64 = out["observables"]
        o["z_vs_fp64_seeds"] = {
            k: float((o[k] - f64[k]["fp64"]) / np.std(f64[k]["fp64_seeds"], ddof=1))
            for k in o if k in f64}
        out["oracle_seed0"] = o
        log(f"oracle: {o}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--part", choices=("trajectory", "statistics", "rotation"), required=True)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--rot-steps", type=int, default=100000)
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--oracle-seeds", type=int, default=1)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    res = dict(trajectory=part_trajectory, statistics=part_statistics,
               rotation=part_rotation)[args.part](args)
    res["seconds"] = time.perf_counter() - T0
    res["oracle_threads"] = THREADS
    text = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text, flush=True)


if __name__ == "__main__":
    main()
