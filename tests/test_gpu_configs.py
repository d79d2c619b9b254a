"""GPU parity at the stated sizes of the benchmark workloads, against the C
oracle's Verlet-list path (pair_mode 1, pinned against the all-pairs path in
tests/test_oracle.py) on the host's cores.

- 28 beads (4 chains x 7 in a 20 x 20 box): the full 10^4 steps in FP64
  with in-kernel Philox noise within 1e-10 L.
- 20k beads (1000 chains x 20, phi 0.3, WCA, kappa 10, chirality 0.2): 100
  SRK2 steps in FP64 with in-kernel Philox noise, which the oracle reproduces
  (orc_bead_noise), within 1e-10 L; the mixed form's error beside it.
- 400k beads (20k chains x 20, phi 0.4, LJ 2.5 sigma, kappa 5, chirality
  U[-0.5, 0.5]): 10 steps in FP64 within 1e-10 L, with the tiling and cell
  list the benchmark runs; the mixed form's error after 1 and 10 steps.
- The 256-replica ensemble (chain counts log-U[50, 2000], lengths U{5..40}):
  the whole pool in FP64 for 100 steps, the smallest, the largest and six
  random replicas against the oracle within 1e-10 L, every replica checked
  for its step count and geometry, mixed against FP64 on every replica; then
  every replica for 10^4 steps in the mixed form.
- A serving stream (Poisson arrivals, bead counts log-U[1k, 200k], ABPs and
  chains of 5-40, 2D and 3D) through the continuous-batching server in FP64
  under an HBM budget that makes requests wait for room: every served result
  against the oracle within 1e-10 L (step counts scaled to 2-200 so the
  oracle finishes in seconds).
"""
import os

import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O
from gcr_amd import engine as E
from gcr_amd import ensemble as ENS
from gcr_amd import serving as SV
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu

POS_TOL = 1e-10  # times L
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def sysd_of(w):
    p = w.phys
    return dict(dim=p.dim, n=p.n_beads, n_nearby=p.n_nearby, has_springs=p.has_springs,
                has_wca=p.has_wca, has_propel=p.has_propel, L=p.L, D=p.D, Dr=p.Dr,
                radius=p.radius, wca=np.asarray(p.wca), spring=np.asarray(p.spring), v0=p.v0)


def oracle(w, steps, tv=None):
    """The oracle's run of w for `steps` steps from its time_vals, with the
    in-kernel noise's counter-based generator (seed = the replica's key)."""
    tv = (w.time_vals if tv is None else tv).copy()
    tv[7] = steps
    v0c = getattr(w, "v0_per_chain", None)
    sysd = sysd_of(w)
    if v0c is not None and np.any(v0c != w.phys.v0):
        raise AssertionError("per-chain speeds differ from phys.v0")
    return O.srk2_run(sysd, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle, tv,
                      xi=None, seed=w.phys.seed, pair_mode=1, n_threads=THREADS)


def err(a, b, L):
    return float(np.max(np.abs(N.min_image(a - b, L))))


def run_single(w, precision, steps, skin=None, tile=0, threads=0, batch=64, rebuild=8):
    phys = w.phys
    if skin is not None:
        import copy
        phys = copy.copy(w.phys)
        phys.skin = skin
    p = E.Pool(phys.dim, precision, bead_capacity=w.n, cell_capacity=E.replica_cells(phys) + 1,
               chain_capacity=phys.n_chains, replica_capacity=1)
    try:
        p.set_schedule(batch, rebuild)
        p.set_tile(tile)
        p.set_tile_threads(threads)
        rid = p.add(phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        tv = w.time_vals.copy()
        tv[7] = steps
        p.begin(rid, tv)
        p.run()
        return p.status(rid), p.download(rid, tangent=False, forces=False)
    finally:
        p.close()


def test_chains_4x7_fp64_full_length():
    """The CPU-runnable workload (4 flexible chains of 7 beads in a 20 x 20
    box, WCA, springs, tangent propulsion, T = 0.01, dt 1e-3) for its full
    10^4 steps in FP64 with the in-kernel noise, against the oracle: within
    1e-10 L after 100, 1000 and 10^4 steps (the error is printed at each;
    about 4e-11 after 10^4 steps, 2e-9 being the bar)."""
    w = W.chains_4x7(seed=1, steps=10000)
    L = w.phys.L
    errs = {}
    for steps in (100, 1000, 10000):
        pos_o, pos1_o, _, _, st_o = oracle(w, steps)
        st, out = run_single(w, E.GCR_FP64, steps)
        assert st["steps"] == int(st_o[2]) == steps
        errs[steps] = err(out["pos"], pos_o, L)
        if steps == 100:
            assert st["elapsed"] == pytest.approx(st_o[1], rel=1e-10)
            assert err(out["pos1"], pos1_o, L) <= POS_TOL * L
    print(f"[4 x 7 chains] fp64 vs oracle: " +
          ", ".join(f"{k} steps {v:.3e}" for k, v in errs.items()) + f" (L = {L:g})", flush=True)
    for steps, e in errs.items():
        assert e <= POS_TOL * L, (steps, e)


def test_chains_1000x20_fp64_100_steps():
    w = W.chains_1000x20(seed=0, steps=100)
    assert w.n == 20000
    pos_o, pos1_o, _, _, st_o = oracle(w, 100)
    st, out = run_single(w, E.GCR_FP64, 100)
    L = w.phys.L
    e64 = err(out["pos"], pos_o, L)
    st_m, out_m = run_single(w, E.GCR_MIXED, 100)
    em = err(out_m["pos"], pos_o, L)
    print(f"[20k beads] fp64 {e64:.3e}  mixed {em:.3e} (L = {L:.2f})", flush=True)
    assert st["steps"] == st_m["steps"] == int(st_o[2]) == 100
    assert e64 <= POS_TOL * L
    assert err(out["pos1"], pos1_o, L) <= POS_TOL * L
    assert st["dt"] == pytest.approx(st_o[0], rel=1e-10)
    assert st["elapsed"] == pytest.approx(st_o[1], rel=1e-10)
    assert em <= 1e-6 * L


def test_chains_20kx20_fp64_10_steps_at_bench_settings():
    w = W.chains_20kx20(seed=1, steps=10)
    assert w.n == 400000
    rc = float(w.phys.wca[3])
    pos_o, pos1_o, _, _, st_o = oracle(w, 10)
    L = w.phys.L
    st, out = run_single(w, E.GCR_FP64, 10, skin=0.3 * rc, tile=8, threads=256, rebuild=32)
    e64 = err(out["pos"], pos_o, L)
    assert st["steps"] == 10 and int(st_o[2]) == 10
    assert e64 <= POS_TOL * L
    assert err(out["pos1"], pos1_o, L) <= POS_TOL * L
    assert st["dt"] == pytest.approx(st_o[0], rel=1e-10)
    assert st["max_forces_strength"] == pytest.approx(st_o[3], rel=1e-10)
    # the mixed form after one step and after ten
    pos1s_o, *_ = oracle(w, 1)
    _, out1 = run_single(w, E.GCR_MIXED, 1, skin=0.3 * rc, tile=8, threads=256, rebuild=32)
    _, out10 = run_single(w, E.GCR_MIXED, 10, skin=0.3 * rc, tile=8, threads=256, rebuild=32)
    e1, e10 = err(out1["pos"], pos1s_o, L), err(out10["pos"], pos_o, L)
    # The LJ law is cut, not shifted, at rc: a pair whose squared distance
    # lies within FP32 rounding of rc^2 adds F(rc) = |coeff(rc^2)| rc (0.044
    # here) in one precision and nothing in the other, which moves a bead by
    # up to dt F(rc) / r per step (half of it in the corrector); the FP32
    # rounding of the other pair terms moves it by about 1e-8.
    p0, p1, rc2 = w.phys.wca[0], w.phys.wca[1], rc * rc
    f_rc = abs(rc2 ** -2 * (p0 - p1 * rc2 ** -1.5)) * rc
    step_bound = float(w.time_vals[1]) * f_rc / w.phys.radius
    print(f"[400k beads] fp64 10 steps {e64:.3e}; mixed 1 step {e1:.3e}, 10 steps {e10:.3e} "
          f"(one cutoff pair per step moves a bead by up to {step_bound:.2e})", flush=True)
    assert e1 <= 2 * step_bound and e10 <= 10 * 2 * step_bound


@pytest.fixture(scope="module")
def ensemble():
    return W.uneven_replicas(seed=0, n_replicas=256, steps=100)


def test_ensemble_256_replicas_fp64_against_oracle(ensemble):
    ws = ensemble
    assert len(ws) == 256
    steps = 100
    res = {}
    for prec in (E.GCR_FP64, E.GCR_MIXED):
        pool = ENS.pool_for(ws, prec)
        try:
            pool.set_schedule(64, 8)
            rids = ENS.add_all(pool, ws)
            for w, r in zip(ws, rids):
                tv = w.time_vals.copy()
                tv[7] = steps
                pool.begin(r, tv)
            pool.run()
            res[prec] = [(pool.status(r), pool.download(r, pos1=False, tangent=False,
                                                        forces=False)["pos"]) for r in rids]
        finally:
            pool.close()
    sizes = np.array([w.n for w in ws])
    rng = np.random.default_rng(7)
    pick = [int(np.argmin(sizes)), int(np.argmax(sizes))]
    pick += [int(i) for i in rng.choice(np.setdiff1d(np.arange(256), pick), 6, replace=False)]
    worst_m = 0.0
    for i, w in enumerate(ws):
        st, pos = res[E.GCR_FP64][i]
        st_m, pos_m = res[E.GCR_MIXED][i]
        L = w.phys.L
        assert st["steps"] == st_m["steps"] == steps and not st["failed"], i
        assert np.all(np.isfinite(pos)) and np.all((pos >= 0) & (pos <= L)), i
        # no chain broke or folded onto itself: 0 < |bond| < 2 a
        m = w.n // w.phys.n_chains
        b = np.linalg.norm(N.min_image(np.diff(pos.reshape(w.phys.n_chains, m, 2), axis=1), L),
                           axis=2)
        assert 0 < np.min(b) and np.max(b) < 2 * w.rest_lengths[0], i
        worst_m = max(worst_m, err(pos_m, pos, L) / L)
    for i in pick:
        w = ws[i]
        pos_o, _, _, _, st_o = oracle(w, steps)
        st, pos = res[E.GCR_FP64][i]
        e = err(pos, pos_o, w.phys.L)
        print(f"[ensemble] replica {i} n={w.n} fp64 vs oracle {e:.3e}", flush=True)
        assert e <= POS_TOL * w.phys.L, i
        assert st["dt"] == pytest.approx(st_o[0], rel=1e-10)
    print(f"[ensemble] mixed vs fp64, worst over 256 replicas: {worst_m:.3e} L", flush=True)
    assert worst_m <= 1e-6


def test_ensemble_256_replicas_1e4_steps_mixed(ensemble):
    ws = ensemble
    steps = 10000
    pool = ENS.pool_for(ws, E.GCR_MIXED)
    try:
        pool.set_schedule(64, 8)
        rids = ENS.add_all(pool, ws)
        for w, r in zip(ws, rids):
            tv = w.time_vals.copy()
            tv[7] = steps
            pool.begin(r, tv)
        pool.run()
        replays = 0
        for w, r in zip(ws, rids):
            st = pool.status(r)
            replays += st["replays"]
            assert st["steps"] == steps and not st["failed"] and not st["running"], w.name
            pos = pool.download(r, pos1=False, tangent=False, forces=False)["pos"]
            L = w.phys.L
            assert np.all(np.isfinite(pos)) and np.all((pos >= 0) & (pos <= L)), w.name
            # no chain broke or folded onto itself: 0 < |bond| < 2 a
            m = w.n // w.phys.n_chains
            b = np.linalg.norm(N.min_image(np.diff(pos.reshape(w.phys.n_chains, m, 2), axis=1),
                                           L), axis=2)
            assert 0 < np.min(b) and np.max(b) < 2 * w.rest_lengths[0], w.name
        print(f"[ensemble] 256 replicas x 10^4 mixed steps, {replays} replays", flush=True)
    finally:
        pool.close()


def test_serving_stream_fp64_against_oracle():
    reqs = SV.make_stream(seed=0, n_requests=16, scale_steps=0.002)
    srv = SV.Server(hbm_budget_bytes=320 << 20, precision=E.GCR_FP64, steps_per_batch=32,
                    rebuild_every=8, max_replicas=16)
    try:
        done = srv.serve(reqs, keep_results=True, realtime=False)
    finally:
        srv.close()
    assert len(done) == len(reqs)
    first_done = min(r.finished for r in done)
    assert any(r.admitted >= first_done for r in done)  # some waited for room
    dims = {r.dim for r in done}
    abps = {r.work.phys.n_chains == r.n for r in done}
    assert dims == {2, 3} and abps == {True, False}
    assert max(r.n for r in done) > 100000
    for r in sorted(done, key=lambda q: q.rid):
        w = r.work
        st = r.result["status"]
        assert st["steps"] == r.steps and not st["failed"], r.rid
        pos_o, _, T_o, _, st_o = oracle(w, r.steps)
        L = w.phys.L
        e = err(r.result["pos"], pos_o, L)
        print(f"[serving] request {r.rid} dim={r.dim} n={r.n} steps={r.steps} "
              f"vs oracle {e:.3e}", flush=True)
        assert e <= POS_TOL * L, r.rid
        assert np.max(np.abs(r.result["tangent"] - T_o)) <= 1e-9, r.rid
