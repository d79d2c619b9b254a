"""Mixed precision (FP32 pair forces, FP64 accumulation) against the FP64 path.

    python -m gcr_amd.validate_mixed [--chains 20000] [--steps 10000] [--every 100]

Runs the 400k-bead system three times from the same initial state: FP64 with
noise seed A (the reference), mixed with seed A (same noise), and FP64 with
seed B (same physics, other noise). Reports the position error of mixed vs
FP64 after 1, 10, 100, ... steps, and time series of the chain-centre MSD,
polar order, nematic order and mean chain angular velocity; the FP64 seed-A
vs seed-B difference of each observable is the statistical error that the
mixed-vs-FP64 difference is compared with.
"""
import argparse
import json
import time

import numpy as np

from . import engine as E
from . import observables as OB
from . import workloads as W


def run(w, precision, seed, steps, every, error_points, ref=None):
    phys = w.phys
    phys.seed = seed
    pool = E.Pool(2, precision, bead_capacity=w.n, cell_capacity=4 * w.n + 1024,
                  chain_capacity=phys.n_chains, replica_capacity=1)
    pool.set_schedule(64, 8)
    rid = pool.add(phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
    L = phys.L
    unw = OB.Unwrapper(w.pos, L)
    tv = w.time_vals.copy()
    series = dict(t=[], msd_com=[], polar=[], nematic=[], director=[], omega_chain=[])
    positions = {}
    done, t_sim = 0, 0.0
    phi_prev = OB.chain_mean_angle(w.tangent, w.chain_id)
    marks = sorted(set([m for m in error_points if m <= steps] +
                       list(range(every, steps + 1, every))))
    for m in marks:
        tv[7] = m - done
        pool.begin(rid, tv)
        pool.run()
        st = pool.status(rid)
        tv[0] = st["dt"]
        t_prev = t_sim
        t_sim += st["elapsed"]
        done = m
        out = pool.download(rid, pos1=False, forces=False)
        if m in error_points:
            positions[m] = out["pos"]
        if m % every == 0:
            unw.add(out["pos"])
            phi = OB.chain_mean_angle(out["tangent"], w.chain_id)
            S, d = OB.nematic_order(phi)
            series["t"].append(t_sim)
            series["msd_com"].append(unw.chain_com_msd(w.chain_id))
            series["polar"].append(OB.polar_order(phi))
            series["nematic"].append(S)
            series["director"].append(d)
            dphi = OB.angle_difference(phi, phi_prev)
            series["omega_chain"].append(float(np.mean(dphi) / (t_sim - t_prev)))
            phi_prev = phi
    pool.close()
    return positions, {k: np.array(v) for k, v in series.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=100)
    args = ap.parse_args()
    w = W.chains_20kx20(seed=1, n_chains=args.chains)
    L = w.phys.L
    pts = [1, 10, 100, 1000, args.steps]
    t0 = time.perf_counter()
    pos64, s64 = run(w, E.GCR_FP64, 11, args.steps, args.every, pts)
    posmx, smx = run(w, E.GCR_MIXED, 11, args.steps, args.every, pts)
    _, s64b = run(w, E.GCR_FP64, 22, args.steps, args.every, pts)
    errs = {}
    for m in pts:
        d = posmx[m] - pos64[m]
        d -= L * np.round(d / L)
        a = np.sqrt(np.sum(d * d, axis=1))
        errs[str(m)] = dict(max=float(a.max()), rms=float(np.sqrt(np.mean(a * a))))
    half = len(s64["t"]) // 2
    obs = {}
    for k in ("msd_com", "polar", "nematic", "omega_chain"):
        x64, xmx, xb = s64[k][half:], smx[k][half:], s64b[k][half:]
        obs[k] = dict(fp64=float(np.mean(x64)), mixed=float(np.mean(xmx)),
                      fp64_other_noise=float(np.mean(xb)),
                      mixed_minus_fp64=float(np.mean(xmx) - np.mean(x64)),
                      noise_to_noise=float(np.mean(xb) - np.mean(x64)))
    print(json.dumps({"workload": "400k-bead config (2D, phi 0.4, LJ 2.5 sigma)", "beads": w.n,
                      "steps": args.steps, "sample_every": args.every,
                      "position_error_mixed_vs_fp64": errs,
                      "observables_second_half_means": obs,
                      "final_t": float(s64["t"][-1]), "seconds": time.perf_counter() - t0}))


if __name__ == "__main__":
    main()
