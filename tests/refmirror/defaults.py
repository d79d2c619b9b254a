"""The reference's flags that parameterise the hot path, with the same names
and default values (defaults.py:12-255), so a `sim` dictionary built here
means the same as one built by the reference's cline.cline()."""
import time

import numpy as np


def set_defaults():
    sim = dict(
        filelabel="most_recent",
        n_threads=3,
        dim=2,
        SRK_order=2,
        Lx=50.0,
        walls="none",
        bead_density=0.5,
        chain_length=10,
        n_chains=-1,
        t_max=-1,
        stepnum_max=-1,
        dtmax=4.0e-3,
        RK_cheat_factor=1.0,
        timestep_safety_factor=0.025,
        n_grid_max=0,
        n_grid_min=1,
        grid_factor=2.0,
        T=1e-3,
        repel_d_max=1.0,
        WCA_epsilon=1.0,
        WCA_shift=0.0,
        LJ_cutoff_factor=1.0,
        n_nearby_beads_dont_repel=4,
        b_FENE=False,
        k=100.0,
        a=0.5,
        kappa=1e1,
        spontaneous_bond_angle=0.0,
        rigid_rod_projection_factor=0.0,
        FENE_tolerance=0.1,
        v0=1.0,
        v_tan_angle=0.0,
        b_use_tangent_as_propulsion_direction=True,
        sliding_partner_maxdist=0.0,
        n_motors=0,
        motor_density=0.0,
        D_motors=0.1,
        b_motors_dont_diffuse_while_pushing=True,
        b_one_motor_per_bead=False,
        k_on=0.9,
        k_off=0.2,
        t_wait_v0=-1.0,
        seed=int(time.time()),
        D=-1.0,
        Dr=-1.0,
        pos_init_noise=1.0,
        phi_init=0.0,
        phi_init_noise=1.0e5,
        a_init_factor=1.0e-8,
        t_init_grow=1.0,
        ticker_every=1000,
        save_state_every=-1.0,
        plot_every=500,
    )
    sim["dt0"] = 0.0001 * sim["dtmax"]
    sim["motor_max_dist"] = sim["a"] / 2
    sim["spontaneous_bond_angle"] = 0.0 * np.pi
    sim["v_tan_angle"] = 0.0 * np.pi
    return sim
