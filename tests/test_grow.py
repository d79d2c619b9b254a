"""Growth-phase initialisation (control.py:153-219, populations.py:530-557).

CPU: the schedule and the initial placement against the reference's rules,
written out by hand. GPU: the growth run on the device against the same
schedule driven through the CPU oracle (no noise while growing, so the two
trajectories are deterministic), and the grown state's geometry."""
import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O
from gcr_amd import engine as E
from gcr_amd import grow as G
from gcr_amd import workloads as W


def final_chains(dim=2, n_chains=4, chain_len=7, L=20.0):
    w = W.chains("g", dim, n_chains, chain_len, L=L, kappa=3.0, T=0.01, dtmax=1e-3, seed=1)
    return w


def test_random_initialization_rules():
    pos, t, cid = G.random_initialization(5, 6, 0.5, 20.0, 2, seed=3, a_init_factor=1e-3)
    assert pos.shape == (30, 2) and np.all((pos >= 0) & (pos < 20.0))
    assert np.allclose(np.linalg.norm(t, axis=1), 1.0)
    for c in range(5):
        h = 6 * c
        d = N.min_image(pos[h + 1:h + 6] - pos[h], 20.0)
        want = -1e-3 * 0.5 * np.arange(1, 6)[:, None] * t[h]
        assert np.allclose(d, want, atol=1e-12)
    # the same seed gives the same chains; the generator is NumPy's legacy one
    pos2, _, _ = G.random_initialization(5, 6, 0.5, 20.0, 2, seed=3, a_init_factor=1e-3)
    assert np.array_equal(pos, pos2)
    rng = np.random.RandomState(3)
    head0 = np.mod(10.0 * (1.0 + (2 * rng.rand(2) - 1)), 20.0)
    assert np.allclose(pos[0], head0)


def test_growth_schedule_values():
    w = final_chains()
    f = w.phys
    f.rod = 0.0
    early, rest_e = G.growth_physics(f, 0.1, 7, 0.5)
    assert early.radius == pytest.approx(f.radius * 0.4)
    assert early.wca[3] == pytest.approx(early.radius)        # repulsive-only cutoff = size
    assert early.rod == 0.0 and early.spring[1] == 0.0 and early.spring[0] == 1.0
    assert early.n_nearby == -1 and early.D == 0.0 and not early.has_propel
    assert rest_e[0] == pytest.approx(0.5 * 1e-8 ** (1 - 0.1 ** 0.1))
    late, rest_l = G.growth_physics(f, 0.6, 7, 0.5)
    assert late.radius == f.radius and late.rod == 0.025 and late.spring[1] == 10.0
    assert rest_l[0] == pytest.approx(0.5 * 1e-8 ** (1 - 0.6 ** 0.1))
    abp, _ = G.growth_physics(f, 0.1, 1, 0.5)
    assert abp.radius == f.radius  # single beads keep their size


def sysd_of(ph):
    return dict(dim=ph.dim, n=ph.n_beads, n_nearby=ph.n_nearby, has_springs=ph.has_springs,
                has_wca=ph.has_wca, has_propel=ph.has_propel, L=ph.L, D=ph.D, Dr=ph.Dr,
                radius=ph.radius, wca=np.asarray(ph.wca), spring=np.asarray(ph.spring),
                v0=ph.v0, rod=ph.rod)


def oracle_grow(final, pos, tangent, cid, rest_length, chain_len, tv, t_init_grow, stepskip,
                max_chunks=1_000_000):
    """The growth schedule of G.grow driven through the CPU oracle; also
    returns the progress of each chunk."""
    t = -t_init_grow
    tv = tv.copy()
    tv[3], tv[4] = 1.0, -1.0
    chunks = 0
    progresses = []
    while t <= 0 and chunks < max_chunks:
        progress = max((t + t_init_grow) / t_init_grow, G.EPSILON)
        progresses.append(progress)
        ph, rest = G.growth_physics(final, progress, chain_len, rest_length)
        tv[7] = stepskip
        pos, _, tangent, _, st = O.srk2_run(sysd_of(ph), pos, tangent, cid, rest,
                                            np.zeros(final.n_chains), tv)
        tv[0] = st[0]
        t += st[1]
        chunks += 1
    return pos, tangent, t, progresses


def test_oracle_growth_is_chaotic_once_the_rod_is_on():
    """Why the device is compared with the oracle chunk by chunk rather than over
    the whole schedule: two oracle runs of the schedule whose progress differs
    by one ulp per chunk (as a different summation order of the elapsed time
    gives) stay within a few ulp of each other while the beads grow, and part
    by more than 1e-5 within four chunks after the rigid-rod projection (factor
    0.025, kappa 10) comes on at progress 1/4. The max-force step control feeds each
    predictor's jump back into the next step size, which amplifies any rounding
    difference there."""
    w = final_chains(dim=2)
    f = w.phys
    pos0, t0, cid = G.random_initialization(4, 7, 0.5, f.L, 2, seed=5)
    tv = w.time_vals.copy()
    tv[3], tv[4] = 1.0, -1.0
    runs = [[pos0.copy(), t0.copy(), tv.copy()] for _ in range(2)]
    t = -0.2
    diffs = []
    for chunk in range(10):
        prog = max((t + 0.2) / 0.2, G.EPSILON)
        for k, run in enumerate(runs):
            ph, rest = G.growth_physics(f, np.nextafter(prog, 1.0) if k else prog, 7, 0.5)
            run[2][7] = 10
            run[0], _, run[1], _, st = O.srk2_run(sysd_of(ph), run[0], run[1], cid, rest,
                                                  np.zeros(4), run[2])
            run[2][0] = st[0]
            if k == 0:
                elapsed = st[1]
        t += elapsed
        diffs.append((ph.rod, np.abs(N.min_image(runs[0][0] - runs[1][0], f.L)).max()))
    assert max(d for rod, d in diffs if rod == 0.0) < 1e-12
    assert max(d for rod, d in diffs if rod != 0.0) > 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_growth_on_device_matches_oracle(dim):
    """The bead-growth part of the schedule (progress < 1/4: sizes and rest
    lengths ramping, no rod), run on the device through set_physics between
    chunks, equals the oracle's run of the same chunks."""
    w = final_chains(dim=dim, L=20.0 if dim == 2 else 9.0)
    f = w.phys
    pos0, t0, cid = G.random_initialization(4, 7, 0.5, f.L, dim, seed=5)
    print(f"[grow parity] dim={dim}", flush=True)
    tv = w.time_vals.copy()
    # 0.2 / 6 chunks of 10 steps reach progress ~0.2 with these settings
    pos_g, tan_g, info = G.grow(f, pos0, t0, cid, 0.5, 7, tv, t_init_grow=0.2,
                                growing_stepskip=10, max_chunks=6)
    pos_o, tan_o, t_o, prog = oracle_grow(f, pos0, t0, cid, 0.5, 7, tv, 0.2, 10, max_chunks=6)
    print("grow:", info, "oracle t", t_o, flush=True)
    assert info["chunks"] == 6 and max(prog) < 0.25
    assert info["t"] == pytest.approx(t_o, rel=1e-10)
    assert np.max(np.abs(N.min_image(pos_g - pos_o, f.L))) <= 1e-9 * f.L
    assert np.allclose(tan_g, tan_o, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_growth_rod_phase_chunk_matches_oracle(dim):
    """One chunk of the rod phase (progress 0.6: full size, rod 0.025, kappa 10,
    rest lengths still short of a) from a state grown by the oracle, with the
    physics switched in place on a replica added in the early-growth physics."""
    w = final_chains(dim=dim, L=20.0 if dim == 2 else 9.0)
    f = w.phys
    pos0, t0, cid = G.random_initialization(4, 7, 0.5, f.L, dim, seed=5)
    tv = w.time_vals.copy()
    pos_s, tan_s, t_s, _ = oracle_grow(f, pos0, t0, cid, 0.5, 7, tv, 0.2, 10, max_chunks=8)
    print(f"[grow rod chunk] dim={dim}", flush=True)
    early, rest_e = G.growth_physics(f, 0.1, 7, 0.5)
    early.skin = 0.3 * f.radius + (f.radius - float(early.wca[3]))
    ph, rest = G.growth_physics(f, 0.6, 7, 0.5)
    pool = E.Pool(dim, E.GCR_FP64, bead_capacity=f.n_beads, cell_capacity=4 * f.n_beads + 1024,
                  chain_capacity=f.n_chains, replica_capacity=1)
    try:
        pool.set_schedule(10, 10)
        rid = pool.add(early, pos_s, tan_s, cid, rest_e, np.zeros(f.n_chains))
        pool.set_physics(rid, ph, rest_lengths=rest)
        tvr = tv.copy()
        tvr[0], tvr[3], tvr[4], tvr[7] = 1e-4, 1.0, -1.0, 10
        pool.begin(rid, tvr)
        pool.run()
        st = pool.status(rid)
        out = pool.download(rid, pos1=False, forces=False)
    finally:
        pool.close()
    assert ph.rod == 0.025 and not st["failed"] and st["steps"] == 10
    pos_o, _, tan_o, _, st_o = O.srk2_run(sysd_of(ph), pos_s, tan_s, cid, rest,
                                          np.zeros(f.n_chains), tvr)
    assert st["elapsed"] == pytest.approx(st_o[1], rel=1e-9)
    assert np.max(np.abs(N.min_image(out["pos"] - pos_o, f.L))) <= 1e-9 * f.L


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_growth_full_schedule_geometry(dim):
    """The whole schedule on the device ends, like the oracle's run, just past
    t = 0 with straight chains whose bonds match the oracle's to a few percent
    (the rod phase is chaotic, so the trajectories themselves differ)."""
    L = 20.0 if dim == 2 else 9.0
    w = final_chains(dim=dim, L=L)
    f = w.phys
    pos0, t0, cid = G.random_initialization(4, 7, 0.5, L, dim, seed=5)
    tv = w.time_vals.copy()
    print(f"[grow full] dim={dim}", flush=True)
    pos_g, tan_g, info = G.grow(f, pos0, t0, cid, 0.5, 7, tv, t_init_grow=0.2,
                                growing_stepskip=10)
    pos_o, _, t_o, _ = oracle_grow(f, pos0, t0, cid, 0.5, 7, tv, 0.2, 10)
    print("grow:", info, "oracle t", t_o, flush=True)
    assert 0.0 < info["t"] <= 10 * tv[1] + 1e-12
    assert np.all(np.isfinite(pos_g)) and np.allclose(np.linalg.norm(tan_g, axis=1), 1.0)

    def geometry(p):
        c = p.reshape(4, 7, dim)
        b = np.linalg.norm(N.min_image(np.diff(c, axis=1), L), axis=2)
        ee = np.linalg.norm(N.min_image(c[:, -1] - c[:, 0], L), axis=1)
        return b, ee / b.sum(axis=1)

    b_g, straight_g = geometry(pos_g)
    b_o, _ = geometry(pos_o)
    assert np.all(straight_g > 0.999)
    assert b_g.mean() == pytest.approx(b_o.mean(), rel=0.03)
    assert b_g.min() > 0.9 * b_o.mean() and b_g.max() < 1.1 * b_o.mean()
