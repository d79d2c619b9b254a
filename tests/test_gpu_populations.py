"""Several interacting populations (updates.py:132-267, SRK.py:287-293) on the
GPU against the CPU oracle's multi-population loop (orc_multi_srk2_run): the
interactions() matrix with averaged WCA laws (populations.py:226-232), a
special interaction (the attractive LJ law of examples/active_and_passive_pops.py),
fixed obstacle populations, 2D and 3D. Positions after 100 SRK2 steps with
injected noise within 1e-10 * L."""
import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O
from gcr_amd import engine as E
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu

POS_TOL = 1e-10


def make_pops(dim, L, specs, seed):
    """Populations placed together without overlap (one gcr_init_chains call
    over every chain), then split. specs: dicts with n_chains, chain_len and
    Physics overrides; returns a list of dicts (phys, pos, tangent, chain_id,
    rest_lengths, v_tan_angle)."""
    lens = np.concatenate([np.full(s["n_chains"], s["chain_len"]) for s in specs])
    pos, tang = E.init_chains(dim, L, lens, 0.5, min_dist=0.9, seed=seed)
    out, b0 = [], 0
    for k, s in enumerate(specs):
        nc, m = s["n_chains"], s["chain_len"]
        n = nc * m
        r = s.get("radius", 1.0)
        wca = W.wca_params(r, 1.0, s.get("shift", 0.0), s.get("cutoff_factor", 1.0))
        T = s.get("T", 0.01)
        D = T / r
        phys = E.Physics(dim=dim, n_beads=n, n_chains=nc, L=L, D=D,
                         Dr=s.get("Dr", 3 * D / (4 * r * r)), radius=r, wca=wca,
                         spring=np.array([100.0, s.get("kappa", 4.0), 0.0, 0.0, float(m > 1), 0.1]),
                         v0=s.get("v0", 1.0), n_nearby=4, has_springs=m > 1, has_wca=True,
                         has_propel=True, seed=seed * 31 + k)
        out.append(dict(phys=phys, pos=pos[b0:b0 + n].copy(), tangent=tang[b0:b0 + n].copy(),
                        chain_id=np.repeat(np.arange(nc, dtype=np.int64), m),
                        rest_lengths=np.full(nc, 0.5),
                        v_tan_angle=np.full(nc, s.get("vtan", 0.0)),
                        fixed=bool(s.get("fixed", False))))
        b0 += n
    return out


def averaged_laws(pops, special=None):
    """interpop_forces for every ordered pair: the special law if given, else
    the mean of the two WCA vectors (populations.py:226-232)."""
    laws = {}
    for a, pa in enumerate(pops):
        for b, pb in enumerate(pops):
            if a != b:
                laws[(a, b)] = 0.5 * (np.asarray(pa["phys"].wca) + np.asarray(pb["phys"].wca))
    for key, law in (special or {}).items():
        laws[key] = np.asarray(law, dtype=np.float64)
    return laws


def sysd_of(p):
    ph = p["phys"]
    return dict(dim=ph.dim, n=ph.n_beads, n_nearby=ph.n_nearby, has_springs=ph.has_springs,
                has_wca=ph.has_wca, has_propel=ph.has_propel, L=ph.L, D=ph.D, Dr=ph.Dr,
                radius=ph.radius, wca=np.asarray(ph.wca), spring=np.asarray(ph.spring),
                v0=ph.v0)


def run_both(pops, laws, steps=100, precision=E.GCR_FP64, extra=None):
    dim = pops[0]["phys"].dim
    rng = np.random.default_rng(500)
    xis = [rng.normal(size=(steps, p["phys"].n_beads, 2 * dim)) for p in pops]
    rc_max = max(float(p["phys"].wca[3]) for p in pops)
    mdps, n_upd = W.grid_constants(0.5, 10, max(rc_max, max(l[3] for l in laws.values())))
    tv = np.array([1e-7, 1e-3, mdps, 1.0, -1.0, n_upd, n_upd, steps])
    orc = [dict(sysd=sysd_of(p), pos=p["pos"], tangent=p["tangent"], chain_id=p["chain_id"],
                rest_lengths=p["rest_lengths"], vtan=p["v_tan_angle"], xi=xis[k],
                fixed=p["fixed"]) for k, p in enumerate(pops)]
    out_o, st_o = O.multi_srk2_run(orc, laws, tv)
    nb = sum(p["phys"].n_beads for p in pops) + (extra.n if extra is not None else 0)
    pool = E.Pool(dim, precision, bead_capacity=nb, cell_capacity=1 << 17, chain_capacity=nb,
                  replica_capacity=len(pops) + 1)
    try:
        pool.set_schedule(32, 8)
        rids = [pool.add(p["phys"], p["pos"], p["tangent"], p["chain_id"], p["rest_lengths"],
                         p["v_tan_angle"]) for p in pops]
        pool.group(rids, laws, [k for k, p in enumerate(pops) if p["fixed"]])
        if extra is not None:
            xr = pool.add(extra.phys, extra.pos, extra.tangent, extra.chain_id,
                          extra.rest_lengths, extra.v_tan_angle)
            tve = extra.time_vals.copy()
            tve[7] = steps
            pool.begin(xr, tve)
        for k, r in enumerate(rids):
            if not pops[k]["fixed"]:
                pool.begin(r, tv, xi=xis[k])
        pool.run()
        got = [pool.download(r) for r in rids]
        sts = [pool.status(r) for r in rids]
    finally:
        pool.close()
    return out_o, st_o, got, sts


def check(pops, out_o, st_o, got, sts):
    L = pops[0]["phys"].L
    for k, p in enumerate(pops):
        if p["fixed"]:
            assert np.array_equal(got[k]["pos"], p["pos"])
            continue
        assert np.max(np.abs(N.min_image(got[k]["pos"] - out_o[k][0], L))) <= POS_TOL * L, k
        assert np.max(np.abs(N.min_image(got[k]["pos1"] - out_o[k][1], L))) <= POS_TOL * L, k
        assert sts[k]["steps"] == int(st_o[2])
        assert sts[k]["dt"] == pytest.approx(st_o[0], rel=1e-10)
        assert sts[k]["elapsed"] == pytest.approx(st_o[1], rel=1e-10)


def cross_contacts(pops, laws):
    """Pairs of different populations within their law's cutoff at the start."""
    L = pops[0]["phys"].L
    n = 0
    for (a, b), law in laws.items():
        d = N.min_image(pops[a]["pos"][:, None, :] - pops[b]["pos"][None, :, :], L)
        n += int(np.sum(np.sum(d * d, axis=-1) <= law[3] ** 2))
    return n


def test_two_chain_populations_averaged_law():
    pops = make_pops(2, 14.0, [dict(n_chains=20, chain_len=6, radius=1.0, vtan=0.2,
                                    cutoff_factor=W.LJ_25_SIGMA),
                               dict(n_chains=30, chain_len=3, radius=0.8, v0=0.0, kappa=0.0)],
                     seed=3)
    laws = averaged_laws(pops)
    print("[populations] averaged 2D", cross_contacts(pops, laws), flush=True)
    assert cross_contacts(pops, laws) > 0
    check(pops, *run_both(pops, laws))


def test_special_attractive_interaction():
    """examples/active_and_passive_pops.py: long self-propelled chains and
    short passive ones sticking through an LJ law cut at 2 r_WCA."""
    pops = make_pops(2, 16.0, [dict(n_chains=12, chain_len=10, vtan=0.1),
                               dict(n_chains=40, chain_len=2, v0=0.0, kappa=0.0)], seed=4)
    sticky = W.wca_params(1.0, 1.0, 0.0, 2.0)
    laws = averaged_laws(pops, {(0, 1): sticky, (1, 0): sticky})
    print("[populations] sticky", cross_contacts(pops, laws), flush=True)
    assert cross_contacts(pops, laws) > 0
    check(pops, *run_both(pops, laws))


def test_fixed_obstacles_and_three_populations():
    """ABPs, chains and a fixed population of obstacle beads: the fixed beads
    never move, the others feel them, one dt for the two moving ones."""
    pops = make_pops(2, 15.0, [dict(n_chains=60, chain_len=1, Dr=0.2),
                               dict(n_chains=15, chain_len=5, vtan=-0.2),
                               dict(n_chains=40, chain_len=1, fixed=True, radius=1.2)], seed=5)
    laws = averaged_laws(pops)
    del laws[(2, 0)], laws[(2, 1)]  # nothing acts on a fixed population
    print("[populations] fixed", cross_contacts(pops, laws), flush=True)
    assert cross_contacts(pops, laws) > 0
    check(pops, *run_both(pops, laws))


def test_populations_3d():
    pops = make_pops(3, 8.0, [dict(n_chains=20, chain_len=4), dict(n_chains=60, chain_len=1)],
                     seed=6)
    laws = averaged_laws(pops)
    print("[populations] 3D", cross_contacts(pops, laws), flush=True)
    assert cross_contacts(pops, laws) > 0
    check(pops, *run_both(pops, laws))


def test_group_beside_an_independent_replica():
    """A group and an unrelated replica in one pool: the group's result is
    that of the oracle, the other replica is untouched by it."""
    pops = make_pops(2, 14.0, [dict(n_chains=20, chain_len=6), dict(n_chains=30, chain_len=1)],
                     seed=7)
    laws = averaged_laws(pops)
    other = W.chains("other", 2, 30, 5, L=14.0, T=0.01, steps=100, seed=8)
    check(pops, *run_both(pops, laws, extra=other))


def test_mixed_precision_group_close_to_fp64():
    pops = make_pops(2, 14.0, [dict(n_chains=20, chain_len=6), dict(n_chains=30, chain_len=3)],
                     seed=9)
    laws = averaged_laws(pops)
    out_o, st_o, got, sts = run_both(pops, laws, steps=20, precision=E.GCR_MIXED)
    L = pops[0]["phys"].L
    for k in range(2):
        assert np.max(np.abs(N.min_image(got[k]["pos"] - out_o[k][0], L))) < 1e-5


def motor_population(dim, L, n, seed, D=0.0, still=True, k_on=0.9, k_off=0.2, imparts=1.0,
                     max_dist=0.5):
    """Motor_Population (populations.py:617-660) as a population of single
    beads with no forces: fixed when D = 0, else diffusing."""
    rng = np.random.default_rng(seed)
    phys = E.Physics(dim=dim, n_beads=n, n_chains=n, L=L, D=D, Dr=0.0, radius=1.0,
                     wca=W.wca_params(1.0, 1.0, 0.0, 1.0), spring=np.zeros(6), v0=0.0,
                     n_nearby=-1, has_springs=False, has_wca=False, has_propel=False,
                     seed=seed * 7 + 1, motor=True, motor_still=still, motor_max_dist=max_dist,
                     motor_imparts=imparts, k_on=k_on, k_off=k_off)
    t = np.zeros((n, dim))
    t[:, 0] = 1.0
    return dict(phys=phys, pos=rng.uniform(0, L, (n, dim)), tangent=t,
                chain_id=np.arange(n, dtype=np.int64), rest_lengths=np.ones(n),
                v_tan_angle=np.zeros(n), fixed=D == 0.0)


def run_motors(pops, laws, steps=100):
    """Beads (population 0) and motors (population 1): the GPU against the
    oracle, positions, where_active of the beads and the motors' bindings."""
    dim = pops[0]["phys"].dim
    rng = np.random.default_rng(501)
    xis = [rng.normal(size=(steps, p["phys"].n_beads, 2 * dim)) for p in pops]
    mdps, n_upd = W.grid_constants(0.5, 10, 1.0)
    tv = np.array([1e-7, 1e-3, mdps, 1.0, -1.0, n_upd, n_upd, steps])
    orc = []
    for k, p in enumerate(pops):
        d = dict(sysd=sysd_of(p), pos=p["pos"], tangent=p["tangent"], chain_id=p["chain_id"],
                 rest_lengths=p["rest_lengths"], vtan=p["v_tan_angle"], xi=xis[k],
                 fixed=p["fixed"], seed=p["phys"].seed)
        if p["phys"].motor:
            ph = p["phys"]
            d["motor"] = dict(max_dist=ph.motor_max_dist, imparts=ph.motor_imparts, k_on=ph.k_on,
                              k_off=ph.k_off, still=ph.motor_still)
        else:
            d["sysd"]["where_active"] = np.zeros(p["phys"].n_beads, dtype=np.int32)
        orc.append(d)
    out_o, st_o = O.multi_srk2_run(orc, laws, tv)
    nb = sum(p["phys"].n_beads for p in pops)
    pool = E.Pool(dim, E.GCR_FP64, bead_capacity=nb, cell_capacity=1 << 17, chain_capacity=nb,
                  replica_capacity=len(pops))
    try:
        pool.set_schedule(32, 8)
        rids = [pool.add(p["phys"], p["pos"], p["tangent"], p["chain_id"], p["rest_lengths"],
                         p["v_tan_angle"]) for p in pops]
        pool.group(rids, laws, [k for k, p in enumerate(pops) if p["fixed"]])
        for k, r in enumerate(rids):
            if not pops[k]["fixed"]:
                pool.begin(r, tv, xi=xis[k])
        pool.run()
        got = [pool.download(r) for r in rids]
        sts = [pool.status(r) for r in rids]
        was = [pool.where_active(r) for r in rids]
    finally:
        pool.close()
    check(pops, out_o, st_o, got, sts)
    assert np.array_equal(was[0], orc[0]["sysd"]["where_active"])
    assert np.array_equal(was[1], orc[1]["motor_bead"])
    return was


def test_fixed_motors_push_chains():
    """Chains with v0 handed to the motors (config.py:95-98) among fixed
    motors: bindings, unbindings and pushes as the oracle's rule."""
    chains = make_pops(2, 14.0, [dict(n_chains=25, chain_len=6, v0=0.0, vtan=0.0)], seed=11)[0]
    motors = motor_population(2, 14.0, 150, seed=12)
    was = run_motors([chains, motors], {(0, 1): "motor"})
    assert np.sum(was[1] >= 0) > 5  # motors are bound at the end


def test_diffusing_motors_3d():
    chains = make_pops(3, 8.0, [dict(n_chains=20, chain_len=4, v0=0.0)], seed=13)[0]
    motors = motor_population(3, 8.0, 200, seed=14, D=0.05, max_dist=0.6)
    was = run_motors([chains, motors], {(0, 1): "motor"})
    assert np.sum(was[1] >= 0) > 3


def test_moving_population_without_pair_force():
    """A population with no forces among its own beads (no "WCA" in its list)
    that the chains act on through their cross law: its grid holds about four
    beads per cell, its beads move farther than half a cell edge within the
    run, so the chains' scans of its cells stay complete only through its
    re-sorts; positions as the oracle's."""
    import ctypes as C
    from gcr_amd import _lib
    pops = make_pops(2, 20.0, [dict(n_chains=20, chain_len=6),
                               dict(n_chains=150, chain_len=1, T=3.0, v0=8.0)], seed=41)
    pops[1]["phys"].has_wca = False
    laws = averaged_laws(pops)
    assert cross_contacts(pops, laws) > 10
    out_o, st_o, got, sts = run_both(pops, laws, steps=300)
    check(pops, out_o, st_o, got, sts)
    nc = round(np.sqrt(_lib.load().gcr_replica_cells(2, C.byref(pops[1]["phys"].desc()))))
    assert nc > 3
    moved = np.sqrt((N.min_image(out_o[1][0] - pops[1]["pos"], 20.0) ** 2).sum(axis=1))
    assert moved.max() > 0.5 * 20.0 / nc
