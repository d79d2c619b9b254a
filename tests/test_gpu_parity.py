"""GPU parity: the HIP path through the C-ABI against the CPU FP64 oracle.

Tolerances (FP64 path): forces within 1e-11 relative to the largest force;
positions after 100 SRK2 steps with injected noise within 1e-10 * L
(min-image difference), step status within 1e-10 relative. The mixed path
(FP32 pair forces, FP64 accumulation) is held to 1e-5 relative force error.
"""
import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O
from gcr_amd import engine as E
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu

POS_TOL = 1e-10  # times L
FORCE_TOL = 1e-11  # times max |F|


def sysd_of(w):
    p = w.phys
    return dict(dim=p.dim, n=p.n_beads, n_nearby=p.n_nearby, has_springs=p.has_springs,
                has_wca=p.has_wca, has_propel=p.has_propel, L=p.L, D=p.D, Dr=p.Dr,
                radius=p.radius, wca=np.asarray(p.wca), spring=np.asarray(p.spring), v0=p.v0,
                rod=p.rod, order=p.srk_order, walls=p.walls, wall_zone=p.wall_zone)


walled = W.walled


def cases():
    out = {}
    out["chains_4x7"] = W.chains_4x7(seed=1, steps=100)
    out["chains_wca_cut"] = W.chains("chains_wca_cut", 2, 100, 20, phi=0.3, kappa=10.0, vtan=0.2,
                                  T=0.01, dtmax=1e-3, steps=100, seed=2)
    rng = np.random.default_rng(3)
    out["chains_lj_cut"] = W.chains("chains_lj_cut", 2, 100, 20, phi=0.4,
                                  cutoff_factor=W.LJ_25_SIGMA, kappa=5.0,
                                  vtan=rng.uniform(-0.5, 0.5, 100), T=0.05, dtmax=5e-4,
                                  steps=100, seed=3)
    out["abp"] = W.chains("abp", 2, 600, 1, phi=0.3, T=0.01, Dr=0.1, steps=100, seed=4)
    out["chains3d"] = W.chains("chains3d", 3, 60, 8, L=12.0, kappa=4.0, T=0.01, steps=100,
                               seed=5)
    out["abp3d"] = W.chains("abp3d", 3, 300, 1, L=10.0, T=0.01, Dr=0.1, steps=100, seed=6)
    w = W.chains("shift_excl", 2, 40, 10, L=22.0, kappa=3.0, theta0=0.2, shift=0.1,
                 n_nearby=-1, T=0.01, steps=100, seed=7)
    # a spontaneous angle makes forces.py:366-373 ill-conditioned on exactly
    # straight chains (the sign of a ~1e-17 cross product picks the side), so
    # this case starts from bent chains
    w.pos = np.mod(w.pos + np.random.default_rng(77).normal(0, 0.03, w.pos.shape), w.phys.L)
    out["shift_excl"] = w
    # WCA_shift and all-same-chain exclusion without the spontaneous angle:
    # that term, kappa (theta + theta0 sign(cross)) / sin(theta), flips sign
    # when an angle passes through straight, so trajectories through it are
    # not reproducible to rounding
    out["shift_excl_traj"] = W.chains("shift_excl_traj", 2, 40, 10, L=22.0, kappa=3.0,
                                      shift=0.1, n_nearby=-1, T=0.01, steps=100, seed=7)
    out["diffusing_chains"] = W.chains("diffusing", 2, 40, 6, L=20.0, T=0.01, Dr=0.05,
                                       tangent_locked=False, vtan=0.3, steps=100, seed=8)
    # FENE springs (forces.py:331-341): straight chains at rest length, so
    # every bond starts inside the FENE range (|stretch| < 0.3 a)
    w = W.chains("fene", 2, 100, 12, phi=0.3, kappa=6.0, vtan=0.2, T=0.01, dtmax=1e-3,
                 steps=100, seed=12)
    w.phys.spring = w.phys.spring.copy()
    w.phys.spring[3] = 1.0
    w.phys.spring[5] = 0.3
    out["fene"] = w
    # rigid-rod projection of the updates (updates.py:347-499), partial in
    # 2D and full in 3D
    w = W.chains("rod2d", 2, 60, 8, phi=0.25, kappa=4.0, vtan=0.3, T=0.01, dtmax=1e-3,
                 steps=100, seed=14)
    w.phys.rod = 0.6
    out["rod2d"] = w
    w = W.chains("rod3d", 3, 40, 6, L=12.0, kappa=4.0, T=0.01, steps=100, seed=15)
    w.phys.rod = 1.0
    out["rod3d"] = w
    # the rod kernel's two layouts at their ends: chains of 80 beads (segments
    # of 64 lanes) and of 3 beads (21 chains per wavefront)
    w = W.chains("rod_long2d", 2, 8, 80, L=200.0, kappa=4.0, vtan=0.1, T=0.01, dtmax=1e-3,
                 steps=100, seed=24)
    w.phys.rod = 0.5
    out["rod_long2d"] = w
    w = W.chains("rod_short3d", 3, 200, 3, L=12.0, kappa=2.0, T=0.01, steps=100, seed=25)
    w.phys.rod = 1.0
    out["rod_short3d"] = w
    # SRK_order 4 (SRK.py:186-321): chains, 3D ABPs, and with the rigid rod
    w = W.chains("srk4_chains", 2, 100, 20, phi=0.3, kappa=10.0, vtan=0.2, T=0.01, dtmax=1e-3,
                 steps=100, seed=16)
    w.phys.srk_order = 4
    out["srk4_chains"] = w
    w = W.chains("srk4_abp3d", 3, 300, 1, L=10.0, T=0.01, Dr=0.1, steps=100, seed=17)
    w.phys.srk_order = 4
    out["srk4_abp3d"] = w
    w = W.chains("srk4_rod2d", 2, 60, 8, phi=0.25, kappa=4.0, vtan=0.3, T=0.01, dtmax=1e-3,
                 steps=100, seed=18)
    w.phys.rod = 0.6
    w.phys.srk_order = 4
    out["srk4_rod2d"] = w
    # WCA walls (forces.py:405-435): ABPs in a 2D box walled in x and y, chains
    # walled in x with WCA_shift, 3D chains walled in z, and SRK4 with walls
    out["walls_abp"] = walled(W.chains("walls_abp", 2, 300, 1, L=20.0, T=0.01, Dr=0.1,
                                       steps=100, seed=19), 3, 0.8)
    out["walls_shift"] = walled(W.chains("walls_shift", 2, 300, 1, L=20.0, shift=0.1,
                                         T=0.01, Dr=0.1, steps=100, seed=20), 1, 0.7)
    out["walls_chains"] = walled(W.chains("walls_chains", 2, 80, 6, L=20.0, kappa=3.0,
                                          vtan=0.2, T=0.01, steps=100, seed=23), 3, 0.7,
                                 margin=0.7)
    out["walls3d"] = walled(W.chains("walls3d", 3, 60, 5, L=10.0, kappa=4.0, T=0.01,
                                     steps=100, seed=21), 4, 0.9, margin=0.7)
    w = walled(W.chains("walls_srk4", 2, 200, 1, L=16.0, T=0.01, Dr=0.1, steps=100,
                        seed=22), 3, 0.8)
    w.phys.srk_order = 4
    out["walls_srk4"] = w
    # a wide skin puts about as many beads within rc + skin as a neighbour
    # list holds: some beads keep a list, the others sum over their cells
    out["list_overflow"] = W.chains("list_overflow", 2, 900, 1, phi=0.55,
                                    cutoff_factor=W.LJ_25_SIGMA, T=0.01, Dr=0.1, steps=100,
                                    seed=9, skin=3.3)
    # 3D with a skin so wide that most beads have more list candidates than a
    # 3D list holds (about 135 within rc + skin = 3.9 against 128): they sum
    # their pairs by the cell scan, the others from LDS (one-cell tiles fit)
    out["list_overflow3d"] = W.chains("list_overflow3d", 3, 1200, 1, L=13.0, T=0.01, Dr=0.1,
                                      steps=100, seed=13, skin=2.9)
    # 3D chains with the 400k-bead system's LJ law at volume fraction 0.15:
    # about 7 beads per cell, so a 4-cell 3D tile with its halo (216 cells)
    # overflows LDS and its beads take the cell scan, while 2- and 3-cell
    # tiles fit (see test_tile_shape_does_not_change_results)
    out["lj3d"] = W.chains("lj3d", 3, 400, 10, phi=0.15, cutoff_factor=W.LJ_25_SIGMA, kappa=5.0,
                           T=0.05, dtmax=5e-4, steps=100, seed=14, skin=0.67)
    # sparse 3D chains (WCA, skin 0.6, about 0.8 beads per 1.6-cell): 16 cells
    # per side, so cubes of 7 and 8 cells (and their 9^3 and 10^3 halos) fit in
    # LDS without a halo axis wrapping onto itself
    out["sparse3d"] = W.chains("sparse3d", 3, 320, 10, L=25.6, kappa=4.0, vtan=0.1, T=0.01,
                               steps=100, seed=15, skin=0.6)
    # about 80 beads per cell: a one-cell tile with its halo fits in LDS, a
    # three-cell one does not (see test_tile_shape_does_not_change_results)
    out["dense_cells"] = W.chains("dense_cells", 2, 1300, 1, phi=0.55,
                                  cutoff_factor=W.LJ_25_SIGMA, T=0.01, Dr=0.1, steps=100,
                                  seed=10, skin=8.0)
    return out


CASES = cases()


def oracle_run(w, xi, tv, seed=None):
    return O.srk2_run(sysd_of(w), w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                      tv, xi=xi, seed=w.phys.seed if seed is None else seed)


def gpu_system(w, precision=E.GCR_FP64):
    return E.System(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                    precision=precision)


def pos_err(a, b, L):
    return np.max(np.abs(N.min_image(a - b, L)))


@pytest.mark.parametrize("name", [c for c in CASES if c != "shift_excl_traj"])
def test_force_eval_parity(name):
    w = CASES[name]
    print(f"[force parity] {name} n={w.n} dim={w.phys.dim}", flush=True)
    F_o, T_o = O.force_eval(sysd_of(w), w.pos, w.chain_id, w.rest_lengths, w.v_tan_angle,
                            w.tangent)
    s = gpu_system(w)
    F_g, T_g = s.eval_forces()
    s.close()
    scale = max(1.0, np.max(np.abs(F_o)))
    assert np.max(np.abs(F_g - F_o)) <= FORCE_TOL * scale
    assert np.max(np.abs(T_g - T_o)) <= 1e-13


@pytest.mark.parametrize("name", [c for c in CASES if c != "shift_excl"])
def test_srk2_100_steps_injected_noise(name):
    w = CASES[name]
    print(f"[srk2 parity] {name} n={w.n}", flush=True)
    rng = np.random.default_rng(100)
    xi = rng.normal(size=(100, w.n, 2 * w.phys.dim))
    tv = w.time_vals.copy()
    tv[7] = 100
    pos_o, pos1_o, T_o, F_o, st_o = oracle_run(w, xi, tv)
    s = gpu_system(w)
    st_g = s.run(tv, xi=xi)
    out = s.download()
    s.close()
    L = w.phys.L
    assert pos_err(out["pos"], pos_o, L) <= POS_TOL * L
    assert pos_err(out["pos1"], pos1_o, L) <= POS_TOL * L
    assert np.max(np.abs(out["tangent"] - T_o)) <= 1e-9
    assert np.max(np.abs(out["forces"] - F_o)) <= 1e-8 * max(1.0, np.max(np.abs(F_o)))
    assert st_g[2] == st_o[2] == 100
    assert st_g[0] == pytest.approx(st_o[0], rel=1e-10)
    assert st_g[1] == pytest.approx(st_o[1], rel=1e-10)
    assert st_g[3] == pytest.approx(st_o[3], rel=1e-10)
    assert np.all((out["pos"] >= 0) & (out["pos"] <= L))


def test_in_kernel_philox_matches_oracle_generator():
    """xi = None on both sides: the GPU's in-kernel Philox/Box-Muller noise and
    the oracle's orc_bead_noise give the same trajectory."""
    w = CASES["chains_wca_cut"]
    tv = w.time_vals.copy()
    tv[7] = 100
    pos_o, _, _, _, st_o = oracle_run(w, None, tv, seed=w.phys.seed)
    s = gpu_system(w)
    s.run(tv)
    out = s.download()
    s.close()
    assert pos_err(out["pos"], pos_o, w.phys.L) <= POS_TOL * w.phys.L


def test_two_calls_equal_one_call():
    """Chunked like the reference driver: the second call starts from the
    first call's dt (status.dt, updates.py:249-256) and the next noise steps."""
    w = CASES["chains_4x7"]
    rng = np.random.default_rng(9)
    xi = rng.normal(size=(100, w.n, 4))
    tv = w.time_vals.copy()
    tv[7] = 100
    pos_o, *_ = oracle_run(w, xi, tv)
    s = gpu_system(w)
    tv1 = tv.copy()
    tv1[7] = 37
    st = s.run(tv1, xi=xi[:37])
    tv2 = tv.copy()
    tv2[0] = st[0]
    tv2[7] = 63
    s.run(tv2, xi=xi[37:])
    out = s.download()
    s.close()
    assert pos_err(out["pos"], pos_o, w.phys.L) <= POS_TOL * w.phys.L


def test_max_dt_usable_clips_time():
    w = CASES["chains_4x7"]
    rng = np.random.default_rng(10)
    xi = rng.normal(size=(50, w.n, 4))
    tv = w.time_vals.copy()
    tv[0] = 1e-3
    tv[7] = 50
    tv[4] = 0.0123  # stops after 13 steps with elapsed == 0.0123
    pos_o, _, _, _, st_o = oracle_run(w, xi, tv)
    s = gpu_system(w)
    st_g = s.run(tv, xi=xi)
    out = s.download()
    s.close()
    assert st_o[2] == st_g[2] == 13
    assert st_g[1] == pytest.approx(0.0123, rel=1e-14)
    assert pos_err(out["pos"], pos_o, w.phys.L) <= POS_TOL * w.phys.L


def test_deterministic_bits():
    w = CASES["chains_lj_cut"]
    tv = w.time_vals.copy()
    tv[7] = 60
    res = []
    for _ in range(2):
        s = gpu_system(w)
        s.run(tv)
        res.append(s.download())
        s.close()
    for k in res[0]:
        assert np.array_equal(res[0][k], res[1][k]), k


@pytest.mark.parametrize("order", [2, 4])
def test_rebuild_period_and_replay(order):
    """Results do not depend on the cell-list schedule beyond rounding; a skin
    too thin for the schedule is caught and the batch replayed (SRK2, and
    SRK4 with its three virtual stages)."""
    w = CASES["chains_wca_cut"]
    rng = np.random.default_rng(11)
    xi = rng.normal(size=(64, w.n, 4))
    tv = w.time_vals.copy()
    tv[7] = 64
    sysd = sysd_of(w)
    sysd["order"] = order
    pos_o, *_ = O.srk2_run(sysd, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                           tv, xi=xi)
    print(f"[replay] order={order}", flush=True)
    outs = []
    for K, skin in ((1, 0.3), (16, 0.3), (32, 0.1)):
        phys = W.chains("c", 2, 100, 20, phi=0.3, kappa=10.0, vtan=0.2, T=0.01, dtmax=1e-3,
                        steps=64, seed=2, skin=skin).phys
        phys.srk_order = order
        p = E.Pool(2, E.GCR_FP64, bead_capacity=4096, cell_capacity=1 << 16,
                   chain_capacity=256, replica_capacity=2)
        p.set_schedule(32, K)
        rid = p.add(phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        p.begin(rid, tv, xi=xi)
        p.run()
        st = p.status(rid)
        outs.append((K, skin, st["replays"], p.download(rid)["pos"]))
        p.close()
    for K, skin, replays, pos in outs:
        assert pos_err(pos, pos_o, w.phys.L) <= POS_TOL * w.phys.L, (K, skin)
    assert outs[2][2] > 0  # the thin skin forced a replay
    assert outs[0][2] == 0 and outs[1][2] == 0


def test_alone_vs_mixed_pool_bit_identical():
    """A replica gives the same bits alone and in a pool with others, SRK2
    and SRK4 replicas mixed (the SRK2 ones skip the inner SRK4 stages)."""
    ws = [W.chains(f"m{i}", 2, 30 + 20 * i, 5 + 3 * i, phi=0.25, vtan=0.1 * i, T=0.01,
                   steps=80, seed=20 + i) for i in range(4)]
    ws[1].phys.srk_order = 4
    ws[3].phys.srk_order = 4

    def run(which):
        p = E.Pool(2, E.GCR_MIXED, bead_capacity=1 << 14, cell_capacity=1 << 16,
                   chain_capacity=1024, replica_capacity=8)
        p.set_schedule(16, 8)
        rids = [p.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
                for w in which]
        for w, r in zip(which, rids):
            tv = w.time_vals.copy()
            tv[7] = 80
            p.begin(r, tv)
        p.run()
        res = [p.download(r) for r in rids]
        p.close()
        return res

    mix = run(ws)
    for i, w in enumerate(ws):
        alone = run([w])[0]
        for k in alone:
            assert np.array_equal(alone[k], mix[i][k]), (i, k)


@pytest.mark.parametrize("precision", [E.GCR_FP64, E.GCR_MIXED])
@pytest.mark.parametrize("name,sides", [("chains_lj_cut", ((1, 256), (2, 512), (5, 1024), (16, 512),
                                                          ((9, 7), 256), ((16, 4), 512))),
                                        ("dense_cells", ((1, 256), (3, 512))),
                                        ("list_overflow", ((1, 512), (4, 256))),
                                        ("chains3d", ((1, 256), (2, 512), (6, 1024))),
                                        ("lj3d", ((2, 256), (3, 512), (4, 256))),
                                        ("list_overflow3d", ((1, 256), (3, 512))),
                                        ("sparse3d", ((8, 512), (7, 256), (5, 512), (3, 1024)))])
def test_tile_shape_does_not_change_results(name, sides, precision):
    """Separations are formed from cell-relative positions in every pair path,
    and lists keep the cell-scan order, so the tile edge and the threads per
    tile (and with them whether
    a bead's pairs come from its LDS list or from the global cell scan: in
    dense_cells a 3-cell tile does not fit in LDS, in lj3d a 4-cell 3D tile)
    leaves every bit of the trajectory unchanged."""
    w = CASES[name]
    print(f"[tile shapes] {name} sides={sides} precision={precision}", flush=True)
    tv = w.time_vals.copy()
    tv[7] = 50
    outs = []
    for side, threads in sides:
        p = E.Pool(w.phys.dim, precision, bead_capacity=w.n,
                   cell_capacity=E.replica_cells(w.phys) + 1,
                   chain_capacity=w.phys.n_chains, replica_capacity=1)
        p.set_schedule(16, 8)
        p.set_tile(side)
        p.set_tile_threads(threads)
        rid = p.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        p.begin(rid, tv)
        p.run()
        outs.append(p.download(rid))
        p.close()
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(o[k], outs[0][k]), k


def test_mixed_precision_force_error():
    w = CASES["chains_lj_cut"]
    F_o, _ = O.force_eval(sysd_of(w), w.pos, w.chain_id, w.rest_lengths, w.v_tan_angle,
                          w.tangent)
    s = gpu_system(w, precision=E.GCR_MIXED)
    F_m, _ = s.eval_forces()
    s.close()
    scale = np.max(np.abs(F_o))
    err = np.max(np.abs(F_m - F_o)) / scale
    print(f"mixed-precision max force error / max force = {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("name", ["fene", "srk4_chains", "walls_chains", "walls_srk4", "rod2d",
                                  "srk4_abp3d", "walls3d"])
def test_mixed_precision_trajectories_of_each_feature(name):
    """The mixed form of the general kernels (SRK4, walls, FENE, rod; 2D tiles
    and 3D) follows the FP64 oracle over 20 steps with injected noise to the
    accuracy of FP32 pair forces."""
    w = CASES[name]
    print(f"[mixed trajectory] {name}", flush=True)
    rng = np.random.default_rng(101)
    xi = rng.normal(size=(20, w.n, 2 * w.phys.dim))
    tv = w.time_vals.copy()
    tv[7] = 20
    pos_o, *_ = oracle_run(w, xi, tv)
    s = gpu_system(w, precision=E.GCR_MIXED)
    s.run(tv, xi=xi)
    out = s.download()
    s.close()
    L = w.phys.L
    assert pos_err(out["pos"], pos_o, L) <= 1e-6 * L


def thin_skin_chains(dim, order):
    """Chains in a box whose cell edge is only 0.004 above the cutoff (L =
    20.08 in 2D, 12.05 in 3D, cutoff 1: 20 and 12 cells per side), so the
    half skin (0.002) is below a
    typical single-step displacement (sqrt(2 D dt) = 0.0045 per axis): most
    predictors leave it (pred_far) and the batches replay in the safe form,
    with a re-sort on the predictor positions before the second stage."""
    if dim == 2:
        w = W.chains(f"thin{dim}", 2, 60, 10, L=20.08, kappa=4.0, vtan=0.2, T=0.01,
                     dtmax=1e-3, steps=64, seed=81, skin=0.004)
    else:
        w = W.chains(f"thin{dim}", 3, 60, 8, L=12.05, kappa=4.0, T=0.01, dtmax=1e-3,
                     steps=64, seed=82, skin=0.004)
    w.phys.srk_order = order
    return w


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("dim", [2, 3])
def test_thin_skin_replays_match_oracle(dim, order):
    """Replays with pred_far re-sorts on X1 (2D tiles and 3D lists) and the
    sticky safe batches that follow two replays in a row: the trajectory
    still matches the oracle. Batches of 8 steps over 64 steps: the first two
    normal batches replay, the rest start in the safe form, so the replica
    sees fewer replays than batches."""
    w = thin_skin_chains(dim, order)
    rng = np.random.default_rng(12)
    xi = rng.normal(size=(64, w.n, 2 * dim))
    tv = w.time_vals.copy()
    tv[7] = 64
    pos_o, pos1_o, *_ = oracle_run(w, xi, tv)
    p = E.Pool(dim, E.GCR_FP64, bead_capacity=w.n, cell_capacity=E.replica_cells(w.phys) + 1,
               chain_capacity=w.phys.n_chains, replica_capacity=1)
    try:
        p.set_schedule(8, 8)
        rid = p.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        p.begin(rid, tv, xi=xi)
        p.run()
        st = p.status(rid)
        out = p.download(rid)
    finally:
        p.close()
    print(f"[thin skin] dim={dim} order={order} replays={st['replays']}", flush=True)
    L = w.phys.L
    assert st["steps"] == 64 and not st["failed"]
    assert 2 <= st["replays"] < 8  # replays, then the sticky safe batches
    assert pos_err(out["pos"], pos_o, L) <= POS_TOL * L
    assert pos_err(out["pos1"], pos1_o, L) <= POS_TOL * L
