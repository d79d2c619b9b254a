"""The drop-in SRK2_force_calc (updates.py:132-267 replacement) driven the way
control.sim_loop drives it (control.py:24-84): population objects with the
reference's interface, a status object and the sim dictionary."""
import numpy as np
import pytest

import gcr_oracle_np as N
import oracle_lib as O
from gcr_amd import dropin
from refmirror import config, defaults, populations
from gcr_amd.engine import init_chains


def make_run(n_chains=6, chain_length=7, Lx=20.0, **over):
    sim = defaults.set_defaults()
    sim.update(dict(n_chains=n_chains, chain_length=chain_length, Lx=Lx, kappa=0.0, T=0.01,
                    dtmax=1e-3, seed=5, t_init_grow=0.0, save_state_every=-1.0))
    sim.update(over)
    sim["dt0"] = 1e-4 * sim["dtmax"]
    config.configure(sim)
    pop = populations.Equal_Size_Chain_Collection(sim)
    pop.random_initialization(seed=3)
    config.grid_list_setup(sim, [pop])
    status = config.status_class(sim, start_t=0.0)
    return sim, pop, status


def test_dropin_rejects_unported_features():
    sim, pop, status = make_run()
    f = dropin.make_SRK2_force_calc(config)
    moving_pop_missing = [pop]
    pop.fixed = True
    with pytest.raises(NotImplementedError, match="no moving"):
        f(moving_pop_missing, status, sim)
    pop.fixed = False
    pop.sliding_partner_maxdist = 1.0
    config.SRK_order = 4
    try:
        with pytest.raises(NotImplementedError, match="active sliding with SRK_order 4"):
            f([pop], status, sim)
    finally:
        config.SRK_order = 2
    pop.sliding_partner_maxdist = 0.0
    own = pop.self_interactions
    pop.self_interactions = lambda: own() + [["Efield", np.zeros(2)]]
    with pytest.raises(NotImplementedError, match="Efield"):
        f([pop], status, sim)
    pop.self_interactions = own
    config.SRK_order = 1
    try:
        with pytest.raises(NotImplementedError, match="SRK_order 2 or 4"):
            f([pop], status, sim)
    finally:
        config.SRK_order = 2


def test_reference_shaped_objects():
    sim, pop, status = make_run()
    assert pop.WCA_params.tolist() == [12.0, 12.0, 0.0, 1.0]
    assert [f[0] for f in pop.self_interactions()] == ["springs", "WCA", "self-propulsion"]
    assert pop.spring_params().tolist() == [100.0, 0.0, 0.0, 0.0, 1.0, 0.1]
    assert pop.D == 0.01 and pop.Dr == 0.0075
    assert config.max_dist_per_step == 0.0125 and config.n_steps_update_grid_list == 40
    assert status.dt == pytest.approx(1e-7)


@pytest.mark.gpu
def test_dropin_chunks_match_oracle():
    """Three driver chunks through the drop-in equal one oracle run with the
    same in-kernel noise; status bookkeeping as updates.py:249-256."""
    sim, pop, status = make_run(v_tan_angle=0.15)
    x0, t0 = pop.pos.copy(), pop.tangent_list.copy()
    f = dropin.make_SRK2_force_calc(config)
    steps_total = 0
    for chunk in (30, 50, 20):
        status.next_output_step = status.stepnum + chunk
        f([pop], status, sim)
        steps_total += chunk
        assert status.stepnum == steps_total
    assert np.all(pop.where_active == -1)
    seed = dropin._mix(sim["seed"], 1)
    sysd = dict(dim=2, n=pop.total_pop, n_nearby=4, has_springs=True, has_wca=True,
                has_propel=True, L=20.0, D=pop.D, Dr=pop.Dr, radius=1.0, wca=pop.WCA_params,
                spring=pop.spring_params(), v0=1.0)
    tv = np.array([sim["dt0"], sim["dtmax"], config.max_dist_per_step, 1.0, -1, 40, 40, 100])
    pos_o, pos1_o, T_o, F_o, st_o = O.srk2_run(sysd, x0, t0, pop.chain_id, pop.rest_lengths,
                                               np.full(6, 0.15), tv, xi=None, seed=seed)
    assert np.max(np.abs(N.min_image(pop.pos - pos_o, 20.0))) < 1e-10 * 20.0
    assert np.max(np.abs(N.min_image(pop.pos1 - pos1_o, 20.0))) < 1e-10 * 20.0
    assert np.allclose(pop.tangent_list, T_o, atol=1e-9)
    assert status.dt == pytest.approx(st_o[0], rel=1e-10)
    # status.t accumulates the elapsed time of every chunk
    assert status.t == pytest.approx(st_o[1], rel=1e-10)


@pytest.mark.gpu
def test_dropin_srk4_matches_oracle():
    """config.SRK_order = 4 through the drop-in equals the oracle's SRK4 run."""
    sim, pop, status = make_run(v_tan_angle=0.15)
    x0, t0 = pop.pos.copy(), pop.tangent_list.copy()
    f = dropin.make_SRK2_force_calc(config)
    config.SRK_order = 4
    try:
        status.next_output_step = 40
        f([pop], status, sim)
    finally:
        config.SRK_order = 2
    seed = dropin._mix(sim["seed"], 1)
    sysd = dict(dim=2, n=pop.total_pop, n_nearby=4, has_springs=True, has_wca=True,
                has_propel=True, L=20.0, D=pop.D, Dr=pop.Dr, radius=1.0, wca=pop.WCA_params,
                spring=pop.spring_params(), v0=1.0, order=4)
    tv = np.array([sim["dt0"], sim["dtmax"], config.max_dist_per_step, 1.0, -1, 40, 40, 40])
    pos_o, *_ = O.srk2_run(sysd, x0, t0, pop.chain_id, pop.rest_lengths, np.full(6, 0.15), tv,
                           xi=None, seed=seed)
    assert np.max(np.abs(N.min_image(pop.pos - pos_o, 20.0))) < 1e-10 * 20.0


def test_dropin_wall_geometry():
    """With sim["walls"] = "xy" the population lists the reference's walls
    force (populations.py:222-223) and the walls sit grid_spacing *
    grid_window / 2 inside the faces (forces.py:417, SRK.py:99-103)."""
    sim, pop, status = make_run(walls="xy")
    try:
        names = [f[0] for f in pop.self_interactions()]
        assert "walls" in names
        wp = dict(pop.self_interactions())["walls"]
        assert wp.tolist() == [12.0, 12.0, 0.0, 1.0, 1.0, 1.0, 0.0]
        assert config.n_grid == 10  # int(20 / (2 * max(2 a, rc)))
    finally:
        config.configure(dict(sim, walls="none"))


@pytest.mark.gpu
def test_dropin_walls_match_oracle():
    """Walls in x and y through the drop-in equal the oracle with the same
    wall placement; beads start inside the walls."""
    sim, pop, status = make_run(walls="xy", v_tan_angle=0.1)
    try:
        zone = 20.0 / config.n_grid * max(1, int(np.ceil(2.0 * 1.0 / (20.0 / config.n_grid)))) / 2
        pop.pos[:] = zone + 1.0 + pop.pos * (20.0 - 2 * zone - 2.0) / 20.0
        x0, t0 = pop.pos.copy(), pop.tangent_list.copy()
        f = dropin.make_SRK2_force_calc(config)
        status.next_output_step = 60
        f([pop], status, sim)
        seed = dropin._mix(sim["seed"], 1)
        sysd = dict(dim=2, n=pop.total_pop, n_nearby=4, has_springs=True, has_wca=True,
                    has_propel=True, L=20.0, D=pop.D, Dr=pop.Dr, radius=1.0,
                    wca=pop.WCA_params, spring=pop.spring_params(), v0=1.0, walls=3,
                    wall_zone=zone)
        tv = np.array([sim["dt0"], sim["dtmax"], config.max_dist_per_step, 1.0, -1, 40, 40, 60])
        pos_o, *_ = O.srk2_run(sysd, x0, t0, pop.chain_id, pop.rest_lengths, np.full(6, 0.1),
                               tv, xi=None, seed=seed)
        assert np.max(np.abs(N.min_image(pop.pos - pos_o, 20.0))) < 1e-10 * 20.0
        assert np.all((pop.pos > zone) & (pop.pos < 20.0 - zone))
    finally:
        config.configure(dict(sim, walls="none"))


def two_populations():
    """examples/active_and_passive_pops.py: self-propelled chains of 10
    ("Kimberly") and passive chains of 2 ("Steve") sticking to each other
    through an LJ law cut at 2 r (a special interaction both ways)."""
    sim = defaults.set_defaults()
    sim.update(dict(n_chains=12, chain_length=10, Lx=20.0, kappa=0.0, T=0.01, dtmax=1e-3,
                    seed=5, t_init_grow=0.0, save_state_every=-1.0))
    sim["dt0"] = 1e-4 * sim["dtmax"]
    config.configure(sim)
    kim = populations.Equal_Size_Chain_Collection(sim, chain_length=10, num_chains=12,
                                                  name="Kimberly")
    steve = populations.Equal_Size_Chain_Collection(sim, chain_length=2, num_chains=40,
                                                    name="Steve")
    steve.v0 = 0
    kim.add_special_interaction("Steve", "WCA", LJ_cutoff_factor=2)
    steve.special_interactions["Kimberly"] = kim.special_interactions["Steve"]
    pos, tang = init_chains(2, 20.0, np.array([10] * 12 + [2] * 40), 0.5, seed=3)
    for pop, sl in ((kim, slice(0, 120)), (steve, slice(120, 200))):
        pop.pos[:] = pos[sl]
        pop.pos1[:] = pos[sl]
        pop.tangent_list[:] = tang[sl]
    config.grid_list_setup(sim, [kim, steve])
    status = config.status_class(sim, start_t=0.0)
    return sim, [kim, steve], status


def test_dropin_population_laws():
    """interpop_forces as the reference forms it: the special law both ways,
    the plain WCA law within each population, nothing else."""
    sim, pops, status = two_populations()
    laws = dropin._cross_laws(pops, [True, True])
    want = populations.make_WCA_params(1.0, 1.0, 0.0, 2.0)
    assert set(laws) == {(0, 1), (1, 0)}
    assert np.allclose(laws[(0, 1)], want) and np.allclose(laws[(1, 0)], want)
    plain = populations.make_WCA_params(1.0, 1.0, 0.0, 1.0)
    assert np.allclose(pops[0].interactions(pops[1])[0][1], plain)  # averaged: equal sizes


@pytest.mark.gpu
def test_dropin_two_populations_match_oracle():
    sim, pops, status = two_populations()
    x0 = [p.pos.copy() for p in pops]
    t0 = [p.tangent_list.copy() for p in pops]
    f = dropin.make_SRK2_force_calc(config)
    status.next_output_step = 50
    f(pops, status, sim)
    assert status.stepnum == 50
    law = populations.make_WCA_params(1.0, 1.0, 0.0, 2.0)
    orc = []
    for k, p in enumerate(pops):
        sysd = dict(dim=2, n=p.total_pop, n_nearby=p.n_nearby_beads_dont_repel,
                    has_springs=True, has_wca=True, has_propel=True, L=20.0, D=p.D, Dr=p.Dr,
                    radius=1.0, wca=p.WCA_params, spring=p.spring_params(), v0=float(p.v0))
        orc.append(dict(sysd=sysd, pos=x0[k], tangent=t0[k], chain_id=p.chain_id,
                        rest_lengths=p.rest_lengths, vtan=np.zeros(p.num_chains()),
                        seed=dropin._mix(sim["seed"], 64 + k)))
    tv = np.array([sim["dt0"], sim["dtmax"], config.max_dist_per_step, 1.0, -1,
                   config.n_steps_update_grid_list, config.n_steps_update_grid_list, 50])
    out, st = O.multi_srk2_run(orc, {(0, 1): law, (1, 0): law}, tv)
    for k, p in enumerate(pops):
        assert np.max(np.abs(N.min_image(p.pos - out[k][0], 20.0))) < 1e-10 * 20.0, k
    assert status.dt == pytest.approx(st[0], rel=1e-10)
    assert status.t == pytest.approx(st[1], rel=1e-10)


@pytest.mark.gpu
def test_dropin_active_sliding_matches_oracle():
    """sliding_partner_maxdist > 0: the population's force list holds
    active_sliding_and_WCA instead of WCA and self-propulsion; the push per
    antiparallel partner follows the reference's split into n_threads chunks,
    and where_active comes back counting the partners."""
    sim, pop, status = make_run(sliding_partner_maxdist=1.5, n_chains=12, chain_length=7)
    assert [f[0] for f in pop.self_interactions()] == ["springs", "active_sliding_and_WCA"]
    x0, t0 = pop.pos.copy(), pop.tangent_list.copy()
    f = dropin.make_SRK2_force_calc(config)
    status.next_output_step = 40
    f([pop], status, sim)
    assert config.n_threads == 3
    seed = dropin._mix(sim["seed"], 1)
    wa = np.zeros(pop.total_pop, dtype=np.int32)
    chunk = dropin.chain_chunks(3, pop.chain_id)[pop.chain_id].astype(np.int32)
    sysd = dict(dim=2, n=pop.total_pop, n_nearby=4, has_springs=True, has_wca=True,
                has_propel=False, L=20.0, D=pop.D, Dr=pop.Dr, radius=1.0, wca=pop.WCA_params,
                spring=pop.spring_params(), v0=0.0, sliding_maxdist=1.5, sliding_v0=1.0,
                chunk=chunk, where_active=wa)
    tv = np.array([sim["dt0"], sim["dtmax"], config.max_dist_per_step, 1.0, -1, 40, 40, 40])
    pos_o, *_ = O.srk2_run(sysd, x0, t0, pop.chain_id, pop.rest_lengths, np.zeros(12), tv,
                           xi=None, seed=seed)
    assert np.max(np.abs(N.min_image(pop.pos - pos_o, 20.0))) < 1e-10 * 20.0
    assert np.array_equal(pop.where_active, wa)


def chains_and_motors(D_motors):
    """BDsim.py:24-25: chains plus a Motor_Population when n_motors > 0; the
    beads' v0 is handed to the motors (config.py:95-98)."""
    sim = defaults.set_defaults()
    sim.update(dict(n_chains=20, chain_length=6, Lx=14.0, kappa=0.0, T=0.01, dtmax=1e-3,
                    seed=6, t_init_grow=0.0, save_state_every=-1.0, n_motors=150,
                    D_motors=D_motors, motor_max_dist=0.5))
    sim["dt0"] = 1e-4 * sim["dtmax"]
    config.configure(sim)
    chains = populations.Equal_Size_Chain_Collection(sim)
    chains.random_initialization(seed=3)
    motors = populations.Motor_Population(sim)
    motors.random_initialization(seed=4)
    config.grid_list_setup(sim, [chains, motors])
    status = config.status_class(sim, start_t=0.0)
    return sim, [chains, motors], status


def test_dropin_motor_laws():
    sim, pops, status = chains_and_motors(0.0)
    assert pops[0].v0 == 0 and sim["motor_imparts_force"] == 1.0
    assert pops[1].fixed and pops[1].self_interactions() == []
    assert dropin._cross_laws(pops, [True, False]) == {(0, 1): "motor"}
    push = dict(pops[0].interactions(pops[1]))["motor_push"]
    assert push.tolist() == [0.25, 0.0, 1.0, 0.9, 0.2]


@pytest.mark.gpu
@pytest.mark.parametrize("D_motors", [0.0, 0.05])
def test_dropin_motors_match_oracle(D_motors):
    sim, pops, status = chains_and_motors(D_motors)
    chains, motors = pops
    x0, t0, m0 = chains.pos.copy(), chains.tangent_list.copy(), motors.pos.copy()
    f = dropin.make_SRK2_force_calc(config)
    status.next_output_step = 40
    f(pops, status, sim)
    assert status.stepnum == 40
    wa = np.zeros(chains.total_pop, dtype=np.int32)
    beads = dict(sysd=dict(dim=2, n=chains.total_pop, n_nearby=4, has_springs=True,
                           has_wca=True, has_propel=True, L=14.0, D=chains.D, Dr=chains.Dr,
                           radius=1.0, wca=chains.WCA_params, spring=chains.spring_params(),
                           v0=0.0, where_active=wa),
                 pos=x0, tangent=t0, chain_id=chains.chain_id, rest_lengths=chains.rest_lengths,
                 vtan=np.zeros(20), seed=dropin._mix(sim["seed"], 64))
    mot = dict(sysd=dict(beads["sysd"], n=150, has_springs=False, has_wca=False,
                         has_propel=False, D=D_motors, Dr=0.0, where_active=None),
               pos=m0, tangent=np.tile([1.0, 0.0], (150, 1)), chain_id=np.arange(150),
               rest_lengths=np.ones(150), vtan=np.zeros(150), fixed=D_motors == 0.0,
               seed=dropin._mix(sim["seed"], 65),
               motor=dict(max_dist=0.5, imparts=1.0, k_on=0.9, k_off=0.2, still=True))
    tv = np.array([sim["dt0"], sim["dtmax"], config.max_dist_per_step, 1.0, -1,
                   config.n_steps_update_grid_list, config.n_steps_update_grid_list, 40])
    out, st = O.multi_srk2_run([beads, mot], {(0, 1): "motor"}, tv)
    assert np.max(np.abs(N.min_image(chains.pos - out[0][0], 14.0))) < 1e-10 * 14.0
    assert np.max(np.abs(N.min_image(motors.pos - out[1][0], 14.0))) < 1e-10 * 14.0
    assert np.array_equal(chains.where_active, wa)
    assert np.array_equal(motors.where_active, mot["motor_bead"])
    assert np.sum(motors.where_active >= 0) > 3


@pytest.mark.gpu
def test_dropin_sees_host_edits():
    """Positions edited on the host between calls are uploaded again."""
    sim, pop, status = make_run()
    f = dropin.make_SRK2_force_calc(config)
    status.next_output_step = 10
    f([pop], status, sim)
    pop.pos[0] = np.mod(pop.pos[0] + 0.01, 20.0)
    edited = pop.pos.copy()
    status.next_output_step = status.stepnum  # zero steps: state must come back unchanged
    f([pop], status, sim)
    assert np.array_equal(pop.pos, edited)
