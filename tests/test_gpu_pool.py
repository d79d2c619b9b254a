"""Ensemble and serving on the GPU: batched runs and the continuous-batching
server give each simulation exactly the result it gets alone."""
import numpy as np
import pytest

from gcr_amd import engine as E
from gcr_amd import ensemble as ENS
from gcr_amd import serving as SV
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_batched_ensemble_runs_every_replica():
    ws = [W.chains(f"e{i}", 2, 20 + 15 * i, 5 + 4 * i, phi=0.1 + 0.08 * i, vtan=0.1 * i - 0.2,
                   v0=0.5 + 0.3 * i, T=1e-3, steps=130, seed=40 + i) for i in range(5)]
    sec, bs, pool, rids = ENS.run_batched(ws, 130, steps_per_batch=32, rebuild_every=8)
    assert bs == sum(w.n for w in ws) * 130
    for w, r in zip(ws, rids):
        st = pool.status(r)
        assert st["steps"] == 130 and not st["running"]
        pos = pool.download(r, pos1=False, tangent=False, forces=False)["pos"]
        assert np.all(np.isfinite(pos)) and np.all((pos >= 0) & (pos <= w.phys.L))
    pool.close()


def small_stream():
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(6):
        dim = 2 if i % 2 == 0 else 3
        abp = i % 3 == 0
        w = SV.make_request_workload(rng, dim, int(rng.integers(200, 1500)), abp, seed=900 + i)
        reqs.append(SV.Request(i, 0.0, w, int(rng.integers(40, 200))))
    return reqs


def test_serving_results_equal_alone():
    reqs = small_stream()
    srv = SV.Server(hbm_budget_bytes=64 << 20, steps_per_batch=32, rebuild_every=8,
                    max_replicas=8)
    done = srv.serve(reqs, keep_results=True, realtime=False)
    srv.close()
    assert len(done) == len(reqs)
    for r in done:
        assert r.result["status"]["steps"] == r.steps
        alone = SV.Request(r.rid, 0.0, r.work, r.steps)
        srv1 = SV.Server(hbm_budget_bytes=64 << 20, steps_per_batch=32, rebuild_every=8,
                         max_replicas=2)
        (a,) = srv1.serve([alone], keep_results=True, realtime=False)
        srv1.close()
        assert np.array_equal(a.result["pos"], r.result["pos"]), r.rid
        assert np.array_equal(a.result["tangent"], r.result["tangent"]), r.rid
    rep = SV.report(done)
    assert rep["requests"] == 6 and rep["p99_latency_s"] >= rep["p50_latency_s"] > 0


def test_autotune_picks_a_measured_setting():
    from gcr_amd import autotune as AT
    w = W.chains("tune", 2, 200, 20, phi=0.3, kappa=10.0, vtan=0.2, T=0.01, dtmax=1e-3,
                 steps=100, seed=2)
    r = AT.tune(w, E.GCR_MIXED, skin_fracs=(0.2, 0.3), rebuilds=(8, 16), tiles=(0, 2, 4),
                threads=(256, 512), steps=64)
    print("autotune:", {k: r[k] for k in ("skin", "rebuild_every", "tile", "threads",
                                           "ms_per_step")}, flush=True)
    # the squares, then rectangles of about the best explicit square's area
    shapes = [t for t in (0, 2, 4)] + [x for t in (2, 4) for x in AT.rectangles(t)]
    assert r["threads"] in (256, 512) and r["tile"] in shapes
    assert r["rebuild_every"] in (8, 16)
    assert any(abs(r["skin"] - f * w.phys.wca[3]) < 1e-12 for f in (0.2, 0.3))
    squares = [t for t in r["tiling_trials"] if not isinstance(t["tile"], tuple)]
    assert len(squares) == 6
    assert all(t["tile"] in shapes for t in r["tiling_trials"])
    assert r["tiling_trials"][0]["tile"] == r["tile"]
    assert r["ms_per_step"] == min(t["ms_per_step"] for t in r["schedule_trials"])


def test_autotune_tunes_3d_tiles():
    from gcr_amd import autotune as AT
    w = W.chains("tune3d", 3, 400, 10, phi=0.15, cutoff_factor=W.LJ_25_SIGMA, kappa=5.0,
                 T=0.05, dtmax=5e-4, steps=100, seed=6)
    r = AT.tune(w, E.GCR_MIXED, skin_fracs=(0.3,), rebuilds=(16,), tiles3=(0, 2, 3),
                threads3=(256, 512), steps=64)
    print("autotune 3D:", {k: r[k] for k in ("tile", "threads", "ms_per_step")}, flush=True)
    assert r["tile"] in (0, 2, 3) and r["threads"] in (256, 512)
    assert len(r["tiling_trials"]) == 6


@pytest.mark.parametrize("kind", ["chains2d", "abp2d", "chains3d"])
def test_device_order_parameters_match_host(kind):
    """gcr_replica_order against the host observables on the downloaded
    tangents (per-chain mean angle, polar and nematic order of
    analysis.py:294-328)."""
    from gcr_amd import observables as OB
    if kind == "chains2d":
        w = W.chains("o2", 2, 300, 10, phi=0.3, kappa=10.0, vtan=0.3, T=0.01, steps=50, seed=31)
    elif kind == "abp2d":
        w = W.chains("oa", 2, 2000, 1, phi=0.3, T=0.01, Dr=0.1, steps=50, seed=32)
    else:
        w = W.chains("o3", 3, 100, 6, L=14.0, kappa=4.0, T=0.01, steps=50, seed=33)
    s = E.System(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                 precision=E.GCR_MIXED)
    tv = w.time_vals.copy()
    tv[7] = 50
    s.run(tv)
    dev = s.order_parameters()
    t = s.download()["tangent"]
    s.close()
    phi = OB.chain_mean_angle(t, w.chain_id)
    S, d = OB.nematic_order(phi)
    print(kind, dev, flush=True)
    assert dev["n_chains"] == w.phys.n_chains
    assert abs(dev["polar"] - OB.polar_order(phi)) < 1e-12
    assert abs(dev["nematic"] - S) < 1e-12
    assert abs(dev["director"] - d) < 1e-9


def test_serving_updates_graphs_in_place():
    """Admitting and retiring requests changes launch sizes, not the kernel
    sequence: after the first capture, the normal batch graph of each lane is
    updated in place, and the safe graph is captured only for a replay."""
    reqs = small_stream()
    srv = SV.Server(hbm_budget_bytes=64 << 20, steps_per_batch=32, rebuild_every=8,
                    max_replicas=8)
    done = srv.serve(reqs, keep_results=True, realtime=False)
    stats = {d: lane.pool.graph_stats() for d, lane in srv.lanes.items()}
    replays = sum(r.result["status"]["replays"] for r in done)
    srv.close()
    print("graph stats", stats, "replays", replays, flush=True)
    for d, st in stats.items():
        assert st["captures"] >= 2, (d, st)
        assert st["updates"] >= 1, (d, st)
        assert st["instantiations"] + st["updates"] == st["captures"]
        if replays == 0:
            assert st["instantiations"] <= 2, (d, st)
