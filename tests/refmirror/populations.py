"""Population objects with the reference's interface, for driving the GPU path.

Mirrors the attributes and the force-list "plugin" interface that the hot
path consumes (populations.py:61-611): state arrays pos, pos1, noise_vec,
where_active, tangent_list, forces; chain layout chain_id, rest_lengths,
phonebook; parameter vectors WCA_params (make_WCA_params), spring_params(),
active_params(); and self_interactions() returning [[name, params], ...].
Constructors take the reference's `sim` dictionary (defaults.set_defaults).

Differences from the reference, by design: random_initialization places
straight chains without overlap (gcr_init_chains) instead of the dot-like
chains the growth phase expands (populations.py:530-557, control.py:153-219);
motors, shells and growth/mitosis are not part of this path. An
optional per-chain chirality array `v_tan_angle_per_chain` extends the
scalar v_tan_angle (a per-population constant reduces to the reference).
"""
import numpy as np

from . import config
from gcr_amd.engine import init_chains


def make_WCA_params(repel_d_max, WCA_epsilon, WCA_shift, LJ_cutoff_factor):
    """populations.py:48-58: [12 eps r^6, 12 eps r^12, shift, cutoff]."""
    s6 = repel_d_max ** 6
    fac = 12 * s6 * WCA_epsilon
    return np.array([fac, fac * s6, WCA_shift, LJ_cutoff_factor * repel_d_max - WCA_shift])


def make_WCA_replacements(pop, kwargs):
    """populations.py:34-45: the population's WCA vector with some of
    repel_d_max, WCA_epsilon, WCA_shift, LJ_cutoff_factor replaced."""
    vals = dict(repel_d_max=pop.repel_d_max, WCA_epsilon=pop.WCA_epsilon,
                WCA_shift=pop.WCA_shift, LJ_cutoff_factor=pop.LJ_cutoff_factor)
    vals.update({k: v for k, v in kwargs.items() if k in vals})
    return make_WCA_params(vals["repel_d_max"], vals["WCA_epsilon"], vals["WCA_shift"],
                           vals["LJ_cutoff_factor"])


class Population:
    """Base population (populations.py:61-141)."""

    def __init__(self, total_pop, sim, name="pop"):
        self.total_pop = total_pop
        self.pos = np.empty((total_pop, config.dim))
        self.pos1 = np.empty_like(self.pos)
        self.noise_vec = np.empty((total_pop, 2 * config.dim))
        self.name = name
        self.special_interactions = {}
        self.v0 = 0.0
        self.where_active = np.full(total_pop, 0)
        self.n_nearby_beads_dont_repel = 0
        self.b_init_growing = False
        self.pos_init_noise = sim["pos_init_noise"]
        self.D = sim["D"]
        self.Dr = sim["Dr"]
        self.initialized = False
        self.fixed = False

    def interactions(self, pop):
        return []

    def self_interactions(self):
        return []

    def add_special_interaction(self, pop_name, force_name, **replacements):
        """populations.py:120-125 (only "WCA" has rules, populations.py:23-31)."""
        if force_name != "WCA":
            raise ValueError("invalid force_name for a special interaction: " + force_name)
        self.special_interactions.setdefault(pop_name, []).append(
            [force_name, make_WCA_replacements(self, replacements)])


class Motor_Population(Population):
    """Motors (populations.py:617-660): points that bind to and push the beads
    of the other populations (their interactions() name "motor_push")."""

    def __init__(self, sim, num_motors=-1, name="motors"):
        n = sim["n_motors"] if num_motors < 0 else num_motors
        super().__init__(n, sim, name=name)
        self.D = sim["D_motors"]
        self.motor_max_dist = sim["motor_max_dist"]
        self.motor_max_dist_sq = self.motor_max_dist ** 2
        self.b_one_motor_per_bead = sim["b_one_motor_per_bead"]
        self.b_motors_dont_diffuse_while_pushing = sim["b_motors_dont_diffuse_while_pushing"]
        self.imparts_force = sim["motor_imparts_force"]
        self.k_on = sim["k_on"]
        self.k_off = sim["k_off"]
        self.where_active = np.full(n, -1)
        if self.D == 0:
            self.fixed = True

    def self_interactions(self):
        if self.D > 0:
            return [["diffuse", np.array([float(self.b_motors_dont_diffuse_while_pushing)])]]
        return []

    def size_scale(self):
        return self.motor_max_dist

    def random_initialization(self, seed=0):
        """populations.py:130-132 (uniform about the box centre), from a seed."""
        rng = np.random.RandomState(seed)
        self.pos[:] = np.mod(config.Lx / 2 * (1.0 + self.pos_init_noise
                                              * (2 * rng.rand(self.total_pop, config.dim) - 1)),
                             config.Lx)
        self.pos1[:] = self.pos
        self.initialized = True


class WCA_Population(Population):
    """Beads with a WCA/LJ pair force (populations.py:143-241)."""

    def __init__(self, total_pop, sim, name="beads"):
        super().__init__(total_pop, sim, name=name)
        self.forces = np.zeros((total_pop, config.dim))
        self.repel_d_max = sim["repel_d_max"]
        self.WCA_epsilon = sim["WCA_epsilon"]
        self.WCA_shift = sim["WCA_shift"]
        self.LJ_cutoff_factor = sim["LJ_cutoff_factor"]
        self.adjust_WCA_params()
        self.sim = sim
        if self.D < 0:  # Stokes-Einstein defaults, populations.py:195-203
            self.D = sim["T"] / self.repel_d_max
        if self.Dr < 0:
            self.Dr = 3 * self.D / (4 * self.repel_d_max ** 2)

    def adjust_WCA_params(self):
        self.WCA_params = make_WCA_params(self.repel_d_max, self.WCA_epsilon, self.WCA_shift,
                                          self.LJ_cutoff_factor)

    def interactions(self, other_pop):
        """populations.py:226-232: with another WCA population, the mean of the
        two WCA vectors."""
        flist = super().interactions(other_pop)
        if isinstance(other_pop, WCA_Population):
            flist = flist + [["WCA", 0.5 * (self.WCA_params + other_pop.WCA_params)]]
        return flist

    def self_interactions(self):
        """populations.py:219-224: WCA, and walls when config.walls has any axis."""
        flist = super().self_interactions() + [["WCA", self.WCA_params]]
        if np.any(config.walls):
            flist.append(["walls", np.append(self.WCA_params, config.walls)])
        return flist


class Oriented_WCA_Population(WCA_Population):
    """Beads with a propulsion direction (populations.py:244-301)."""

    def __init__(self, total_pop, sim, name="beads"):
        super().__init__(total_pop, sim, name=name)
        self.v0 = sim["v0"]
        self.tangent_list = np.empty((total_pop, config.dim))
        self.tangent_list[:] = 0.0
        self.tangent_list[:, 0] = 1.0

    def interactions(self, pop):
        """populations.py:276-287: motors push these beads."""
        flist = super().interactions(pop)
        if isinstance(pop, Motor_Population):
            flist = flist + [["motor_push", np.array([
                pop.motor_max_dist_sq, float(pop.b_one_motor_per_bead), pop.imparts_force,
                pop.k_on, pop.k_off])]]
        return flist


class Chain_Collection(Oriented_WCA_Population):
    """Beads strung into chains (populations.py:305-591)."""

    def __init__(self, num_subpops, total_pop, sim, name="beads", chain_id=None):
        super().__init__(total_pop, sim, name=name)
        self.num_subpops = num_subpops
        if chain_id is None:
            chain_id = np.floor(np.arange(total_pop) / (total_pop / num_subpops)).astype(np.int64)
        self.chain_id = np.asarray(chain_id, dtype=np.int64)
        self.k = sim["k"]
        self.reference_rest_length = sim["a"]
        self.rest_lengths = np.full(num_subpops, sim["a"], dtype=np.float64)
        self.spontaneous_bond_angle = sim["spontaneous_bond_angle"]
        if sim["kappa"] >= 0:
            self.rigid_rod_projection_factor = sim["rigid_rod_projection_factor"]
            self.kappa = sim["kappa"]
        else:
            self.rigid_rod_projection_factor = 1
            self.kappa = 0
        self.n_nearby_beads_dont_repel = sim["n_nearby_beads_dont_repel"]
        self.v_tan_angle = sim["v_tan_angle"]
        self.b_use_tangent_as_propulsion_direction = sim["b_use_tangent_as_propulsion_direction"]
        self.sliding_partner_maxdist = sim["sliding_partner_maxdist"]
        self.b_FENE = sim["b_FENE"]
        self.FENE_tolerance = sim["FENE_tolerance"]
        self.longest_chain_length = int(np.max(np.bincount(self.chain_id)))
        self.b_get_orientations = False
        self.v_tan_angle_per_chain = None

    def spring_params(self):
        return np.array([self.k, self.kappa, self.spontaneous_bond_angle, float(int(self.b_FENE)),
                         float(self.b_get_orientations), self.FENE_tolerance])

    def active_params(self):
        return np.array([self.v0, self.sliding_partner_maxdist, self.v_tan_angle])

    def num_chains(self):
        return self.num_subpops

    def self_interactions(self):
        """populations.py:570-591: springs first (they update tangents), WCA,
        then self-propulsion."""
        flist = super().self_interactions()
        flist.append(["self-propulsion", self.active_params()])
        if self.longest_chain_length > 1:
            flist.insert(0, ["springs", self.spring_params()])
        if self.sliding_partner_maxdist > 0 and self.v0 > 0:
            flist = [f for f in flist if f[0] not in ("WCA", "self-propulsion")]
            flist.append(["active_sliding_and_WCA",
                          np.append(self.WCA_params, self.active_params())])
        return flist

    def random_initialization(self, seed=0, min_dist=0.9):
        lens = np.bincount(self.chain_id).astype(np.int32)
        pos, tang = init_chains(config.dim, config.Lx, lens, self.reference_rest_length,
                                min_dist=min_dist, seed=seed)
        self.pos[:] = pos
        self.pos1[:] = pos
        self.tangent_list[:] = tang
        self.initialized = True


class Equal_Size_Chain_Collection(Chain_Collection):
    """Chains of equal length (populations.py:593-614)."""

    def __init__(self, sim, num_chains=-1, chain_length=-1, name="beads"):
        self.num_subpops = sim["n_chains"] if num_chains < 0 else num_chains
        self.subpop_size = sim["chain_length"] if chain_length < 0 else chain_length
        total = self.num_subpops * self.subpop_size
        chain_id = np.floor(np.arange(total) / self.subpop_size).astype(np.int64)
        super().__init__(self.num_subpops, total, sim, name=name, chain_id=chain_id)
        self.b_get_orientations = (self.subpop_size > 1
                                   and self.b_use_tangent_as_propulsion_direction)

    def chain_length(self):
        return self.subpop_size
