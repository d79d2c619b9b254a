"""Chain initialisation by the reference's growth protocol, run on the device.

The reference starts a run of chains (`b_init_growing`, the default path of
BDsim.py) by placing every chain as a point and growing it:

- random_initialization (populations.py:530-557): chain heads from the legacy
  NumPy generator, `Lx/2 (1 + noise (2u - 1))` per axis, a random direction
  (`pi (u > 1/2) + phi_init + 2 pi phi_init_noise u`, and `acos(2u - 1)` as the
  polar angle in 3D); the other beads behind the head at a_init_factor times
  the rest length;
- set_init_grow (populations.py:234-236, 294-297, 376-389): no same-chain
  repulsion at all, kappa 0, k 1, no WCA shift, no noise, no propulsion;
- per chunk of `growing_stepskip` steps, from t = -t_init_grow to 0
  (control.py:153-219, early_time_procedures): progress
  p = max((t - t_start) / t_init_grow, epsilon); for chains of more than one
  bead the bead size grows from 0 to repel_d_max over the first quarter,
  after which the rigid-rod projection is on (factor >= 0.025) with kappa
  10; rest lengths grow as a * a_init_factor^(1 - p^0.1); the pair law is
  made purely repulsive (updates.py:304-310); RK_cheat_factor 1;
- end_init_grow (populations.py:238-241, 391-399) restores everything.

grow() drives this schedule on one replica of a device pool, changing the
physics in place between chunks (gcr_replica_set_physics), so the
initialisation runs on the GPU like the rest of the hot path. The schedule
itself is host logic (a few scalars per chunk), as in the reference.
"""
import copy
import math

import numpy as np

from . import engine as E
from . import workloads as W

EPSILON = 1e-8  # config.py:78


def random_initialization(n_chains, chain_len, rest_length, L, dim, seed, pos_init_noise=1.0,
                          phi_init=0.0, phi_init_noise=1e5, a_init_factor=1e-8):
    """Positions, tangents and chain ids as populations.random_initialization
    leaves them, drawing from np.random.seed(seed)'s legacy generator in the
    reference's order (head position, flip, angle[, polar angle]) per chain."""
    rng = np.random.RandomState(seed)
    n = n_chains * chain_len
    pos = np.zeros((n, dim))
    tangent = np.zeros((n, dim))
    for c in range(n_chains):
        h = c * chain_len
        head = np.mod(L / 2 * (1.0 + pos_init_noise * (2 * rng.rand(dim) - 1)), L)
        phi = np.mod(math.pi * (rng.rand() > 0.5) + phi_init + 2 * math.pi * phi_init_noise * rng.rand()
                     + math.pi, 2 * math.pi) - math.pi
        if dim == 2:
            t = np.array([math.cos(phi), math.sin(phi)])
        else:
            th = math.acos(2 * rng.rand() - 1)
            t = np.array([math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi), math.cos(th)])
        for j in range(chain_len):
            pos[h + j] = head - a_init_factor * rest_length * t * j
            tangent[h + j] = t
    return np.mod(pos, L), tangent, np.repeat(np.arange(n_chains), chain_len).astype(np.int64)


def growth_physics(final, progress, chain_len, rest_length, a_init_factor=1e-8, eps=1.0,
                   lj_cutoff_factor=1.0):
    """(physics, rest lengths) for one chunk of the growth phase at `progress`
    (control.py:178-215 with the set_init_grow values); `final` is the physics
    after growth (its radius is repel_d_max, its spring vector and rod factor
    the saved values)."""
    ph = copy.copy(final)
    ph.spring = np.array(final.spring, dtype=np.float64).copy()
    repel_save = final.radius
    kappa = 0.0
    rod = 0.0
    if chain_len == 1:
        repel = repel_save
    else:
        repel = repel_save
        frac = min(0.25, progress) / 0.25
        if progress < 0.25:
            repel = repel_save * frac
        elif progress < 1.0:
            rod = max(0.025, final.rod)
            kappa = 10.0
        else:
            kappa = final.spring[1]
    wca = W.wca_params(repel, eps, 0.0, lj_cutoff_factor)
    if wca[0] != 0:
        wca[3] = (wca[1] / wca[0]) ** (1 / 6)  # updates.py:304-310: repulsive only
    ph.wca = wca
    ph.radius = repel
    ph.spring[0] = 1.0
    ph.spring[1] = kappa
    ph.spring[2] = final.spring[2]
    ph.rod = rod
    ph.n_nearby = -1
    ph.D = 0.0
    ph.has_propel = False
    rest = np.full(final.n_chains, rest_length * a_init_factor ** (1 - progress ** 0.1))
    return ph, rest


def growth_schedule(final, chain_len, rest_length, t_init_grow=1.0, a_init_factor=1e-8, eps=1.0,
                    lj_cutoff_factor=1.0, max_chunks=1_000_000):
    """The chunks of the growth phase (control.py:153-219): yields the
    (physics, rest lengths) of each chunk while t <= 0, from t = -t_init_grow;
    the caller sends back the time each chunk advanced."""
    t_start = -t_init_grow
    t = t_start
    chunks = 0
    while t <= 0 and chunks < max_chunks:
        progress = max((t - t_start) / t_init_grow, EPSILON)
        elapsed = yield growth_physics(final, progress, chain_len, rest_length, a_init_factor,
                                       eps, lj_cutoff_factor)
        t += elapsed
        chunks += 1


def grow(final, pos, tangent, chain_id, rest_length, chain_len, time_vals, t_init_grow=1.0,
         growing_stepskip=10, a_init_factor=1e-8, precision=E.GCR_FP64, eps=1.0,
         lj_cutoff_factor=1.0, max_chunks=1_000_000):
    """Run the growth phase on the device from the chains of
    random_initialization. time_vals: the run's [dt0, dtmax, max_dist_per_step,
    ...] (RK_cheat_factor is 1 while growing). Returns (pos, tangent, info)."""
    n = final.n_beads
    # cell grid for the largest cutoff of the schedule (the bead size itself)
    ph0, rest0 = growth_physics(final, EPSILON, chain_len, rest_length, a_init_factor, eps,
                                lj_cutoff_factor)
    rc_max = final.radius
    ph0.skin = 0.3 * rc_max + (rc_max - float(ph0.wca[3]))
    pool = E.Pool(final.dim, precision, bead_capacity=n, cell_capacity=4 * n + 1024,
                  chain_capacity=final.n_chains, replica_capacity=1)
    pool.set_schedule(growing_stepskip, growing_stepskip)
    rid = pool.add(ph0, pos, tangent, chain_id, rest0, np.zeros(final.n_chains))
    t = -t_init_grow
    tv = np.array(time_vals, dtype=np.float64).copy()
    tv[3] = 1.0    # RK_cheat_factor while growing
    tv[4] = -1.0   # no save-time clipping
    chunks = steps = 0
    sched = growth_schedule(final, chain_len, rest_length, t_init_grow, a_init_factor, eps,
                            lj_cutoff_factor, max_chunks)
    try:
        step = next(sched, None)
        while step is not None:
            ph, rest = step
            pool.set_physics(rid, ph, rest_lengths=rest)
            tv[7] = growing_stepskip
            pool.begin(rid, tv)
            pool.run()
            st = pool.status(rid)
            if st["failed"]:
                raise RuntimeError("growth: a bead moved more than half a skin in one step")
            tv[0] = st["dt"]
            t += st["elapsed"]
            steps += st["steps"]
            chunks += 1
            try:
                step = sched.send(st["elapsed"])
            except StopIteration:
                step = None
        out = pool.download(rid, pos1=False, forces=False)
    finally:
        pool.close()
    return out["pos"], out["tangent"], dict(t=t, steps=steps, chunks=chunks, dt=tv[0])
