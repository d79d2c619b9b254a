"""Active sliding (the "active_sliding_and_WCA" force, forces.py:79-83, 114-120)
on the GPU against the CPU oracle: v0 t per antiparallel partner within the
sliding range, partners in the bead's own thread chunk counted from the later
bead only, where_active counting them (-1 + partners of both stages). 100 SRK2
steps with injected noise within 1e-10 * L; where_active exactly."""
import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O
from gcr_amd import dropin
from gcr_amd import engine as E
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu

POS_TOL = 1e-10


def sliding(w, maxdist=1.5, v0=1.0, threads=3):
    """Self-propulsion replaced by active sliding, as populations.py:580-591
    does when sliding_partner_maxdist > 0; chunks of `threads` threads."""
    w.phys.has_propel = False
    w.phys.sliding_maxdist = maxdist
    w.phys.sliding_v0 = v0
    w.chunks = dropin.chain_chunks(threads, w.chain_id)
    return w


def oracle_sys(w):
    p = w.phys
    wa = np.zeros(w.n, dtype=np.int32)
    return dict(dim=p.dim, n=p.n_beads, n_nearby=p.n_nearby, has_springs=p.has_springs,
                has_wca=p.has_wca, has_propel=p.has_propel, L=p.L, D=p.D, Dr=p.Dr,
                radius=p.radius, wca=np.asarray(p.wca), spring=np.asarray(p.spring), v0=p.v0,
                rod=p.rod, sliding_maxdist=p.sliding_maxdist, sliding_v0=p.sliding_v0,
                chunk=np.ascontiguousarray(w.chunks[w.chain_id], dtype=np.int32),
                where_active=wa), wa


def compare(w, steps=100, schedule=None):
    rng = np.random.default_rng(700)
    xi = rng.normal(size=(steps, w.n, 2 * w.phys.dim))
    tv = w.time_vals.copy()
    tv[7] = steps
    sysd, wa_o = oracle_sys(w)
    pos_o, pos1_o, T_o, F_o, st_o = O.srk2_run(sysd, w.pos, w.tangent, w.chain_id,
                                               w.rest_lengths, w.v_tan_angle, tv, xi=xi)
    pool = E.Pool(w.phys.dim, E.GCR_FP64, bead_capacity=w.n, cell_capacity=1 << 17,
                  chain_capacity=w.phys.n_chains, replica_capacity=1)
    try:
        if schedule:
            pool.set_schedule(*schedule)
        rid = pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        pool.set_chunks(rid, w.chunks)
        pool.begin(rid, tv, xi=xi)
        pool.run()
        out = pool.download(rid)
        st = pool.status(rid)
        wa = pool.where_active(rid)
    finally:
        pool.close()
    L = w.phys.L
    assert np.max(np.abs(N.min_image(out["pos"] - pos_o, L))) <= POS_TOL * L
    assert np.max(np.abs(N.min_image(out["pos1"] - pos1_o, L))) <= POS_TOL * L
    assert np.array_equal(wa, wa_o)
    assert st["steps"] == steps
    assert st["dt"] == pytest.approx(st_o[0], rel=1e-10)
    assert st["elapsed"] == pytest.approx(st_o[1], rel=1e-10)
    return wa, st


def test_sliding_chains_2d():
    w = sliding(W.chains("slide2d", 2, 120, 8, phi=0.35, kappa=6.0, T=0.01, steps=100, seed=30))
    print(f"[sliding] 2D chains n={w.n}", flush=True)
    wa, _ = compare(w)
    assert np.sum(wa > 0) > 20  # the pushes happen


def test_sliding_within_the_pair_cutoff():
    """LJ 2.5 sigma pairs: the sliding range lies within the cutoff, so the
    2D partners come from the neighbour lists (the other cases scan cells)."""
    w = sliding(W.chains("slide_lj", 2, 150, 8, phi=0.35, cutoff_factor=W.LJ_25_SIGMA,
                         kappa=6.0, T=0.01, steps=100, seed=35))
    print(f"[sliding] LJ chains n={w.n}", flush=True)
    wa, _ = compare(w)
    assert np.sum(wa > 0) > 20


def test_sliding_abps_one_chunk():
    w = sliding(W.chains("slide_abp", 2, 500, 1, phi=0.4, T=0.01, Dr=0.2, steps=100, seed=31),
                maxdist=1.3, threads=1)
    print(f"[sliding] ABPs n={w.n}", flush=True)
    wa, _ = compare(w)
    assert np.sum(wa > 0) > 20


def test_sliding_chains_3d():
    w = sliding(W.chains("slide3d", 3, 60, 6, L=10.0, kappa=4.0, T=0.01, steps=100, seed=32),
                maxdist=1.4, threads=2)
    print(f"[sliding] 3D chains n={w.n}", flush=True)
    wa, _ = compare(w)
    assert np.sum(wa > 0) > 5


def test_sliding_through_replays():
    """A skin too thin for the re-sort period: replayed batches, same result."""
    w = sliding(W.chains("slide_replay", 2, 120, 8, phi=0.35, kappa=6.0, T=0.01, steps=64,
                         seed=33, skin=0.05))
    print(f"[sliding] replay n={w.n}", flush=True)
    compare(w, steps=64, schedule=(32, 32))


def test_sliding_mixed_precision_follows_fp64():
    w = sliding(W.chains("slide_mixed", 2, 120, 8, phi=0.35, kappa=6.0, T=0.01, steps=20,
                         seed=34))
    rng = np.random.default_rng(702)
    xi = rng.normal(size=(20, w.n, 4))
    tv = w.time_vals.copy()
    tv[7] = 20
    sysd, _ = oracle_sys(w)
    pos_o, *_ = O.srk2_run(sysd, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                           tv, xi=xi)
    pool = E.Pool(2, E.GCR_MIXED, bead_capacity=w.n, cell_capacity=1 << 17,
                  chain_capacity=w.phys.n_chains, replica_capacity=1)
    try:
        rid = pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        pool.set_chunks(rid, w.chunks)
        pool.begin(rid, tv, xi=xi)
        pool.run()
        out = pool.download(rid)
    finally:
        pool.close()
    assert np.max(np.abs(N.min_image(out["pos"] - pos_o, w.phys.L))) <= 1e-6 * w.phys.L
