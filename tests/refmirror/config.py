"""Module globals and time-keeping with the reference's meaning (config.py:13-200).

Like the reference, run-time state lives in module globals (dim, Lx,
SRK_order, n_steps_update_grid_list, max_dist_per_step, RK_cheat_factor,
grid_factor, walls) that configure() and grid_list_setup() fill from `sim`.
"""
import numpy as np

dim = 2
Lx = 1.0
SRK_order = 2
n_steps_update_grid_list = 1
RK_cheat_factor = 1
max_dist_per_step = 1
grid_factor = 2.0
walls = np.array([False, False, False])
n_grid = 1
n_threads = 1
pop_name_list = []


class status_class:
    """General time-keeping object (config.py:37-72)."""

    def __init__(self, sim, start_t=None, start_stepnum=0, start_filenumber=0):
        self.t_start = -sim["t_init_grow"] if start_t is None else start_t
        self.stepnum = start_stepnum
        self.filenumber = start_filenumber
        self.b_calc_active_forces = True
        self.t = self.t_start
        self.dt = sim["dt0"]
        self.max_forces_strength = 0
        self.dt_used = 0
        self.next_save_time = self.t_start
        self.next_output_step = 0
        self.update_grid_list_counter = 0
        self.dtmax = sim["dtmax"]


def configure(sim):
    """config.py:75-119 (derived counts, globals, seed)."""
    global dim, SRK_order, Lx, walls, n_threads
    sim["chain_density"] = sim["bead_density"] / (
        sim["repel_d_max"] * (sim["a"] * (sim["chain_length"] - 1) + sim["repel_d_max"]))
    if sim["n_chains"] < 0:
        sim["n_chains"] = max(1, round(sim["Lx"] * sim["Lx"] * sim["chain_density"]))
    sim["n_beads"] = sim["chain_length"] * sim["n_chains"]
    dim = sim["dim"]
    SRK_order = sim["SRK_order"]
    Lx = sim["Lx"]
    walls = np.array([(s in sim["walls"] or sim["walls"] in ["all", "True"]) for s in "xyz"])
    n_threads = sim["n_threads"]  # cline.py:72
    # config.py:88-100: with motors, the beads' propulsion speed is the motors' push
    if sim.get("motor_density", 0) > 0:
        sim["n_motors"] = int(round(sim["Lx"] * sim["Lx"] * sim["motor_density"]))
    if sim.get("n_motors", 0) > 0:
        sim["motor_imparts_force"] = sim["v0"]
        sim["v0"] = 0
    else:
        sim["motor_imparts_force"] = 0
    np.random.seed(sim["seed"] % (2 ** 32))


def grid_list_setup(sim, pops):
    """config.py:147-200: list-rebuild period and the per-step displacement
    cap that the adaptive dt enforces."""
    global n_grid, grid_factor, n_steps_update_grid_list, max_dist_per_step, RK_cheat_factor
    dists = []
    for pop in pops:
        d = 0.0
        if getattr(pop, "longest_chain_length", 0) > 1:
            d = max(d, 2 * max(pop.rest_lengths))
        if hasattr(pop, "WCA_params"):
            d = max(d, pop.WCA_params[-1], getattr(pop, "sliding_partner_maxdist", 0.0))
        if hasattr(pop, "motor_max_dist"):
            d = max(d, pop.motor_max_dist)
        dists.append(d)
    sim["biggest_interaction_distance"] = max(dists)
    sim["n_grid"] = max(sim["n_grid_min"],
                        int(sim["Lx"] / (sim["grid_factor"] * sim["biggest_interaction_distance"])))
    if sim["n_grid_max"] > 0:
        sim["n_grid"] = min(sim["n_grid"], sim["n_grid_max"])
    sim["grid_spacing"] = sim["Lx"] / sim["n_grid"]
    relevant = [d for d in dists if d < sim["Lx"] / 2]
    sim["n_steps_update_grid_list"] = max(
        1, int(1.0 / (sim["timestep_safety_factor"] * sim["RK_cheat_factor"])))
    sim["max_dist_per_step"] = (0.5 * (sim["grid_factor"] - 1) * min(relevant)
                                / sim["n_steps_update_grid_list"])
    sim["grid_window"] = int(np.ceil(sim["biggest_interaction_distance"] / sim["grid_spacing"]))
    n_grid = sim["n_grid"]
    grid_factor = sim["grid_factor"]
    n_steps_update_grid_list = sim["n_steps_update_grid_list"]
    max_dist_per_step = sim["max_dist_per_step"]
    RK_cheat_factor = sim["RK_cheat_factor"]
