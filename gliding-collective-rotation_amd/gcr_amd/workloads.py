"""Synthetic workloads: the benchmark configurations as concrete inputs.

Parameters follow the reference's flag semantics (defaults.py:12-255):
D = T / repel_d_max and Dr = 3 D / (4 r^2) when left negative
(populations.py:195-203), WCA parameters from make_WCA_params
(populations.py:48-58), max_dist_per_step and the list-rebuild period from
config.grid_list_setup (config.py:147-193), dt0 = 1e-4 dtmax (defaults.py:59).
"""
from dataclasses import dataclass

import numpy as np

from .engine import Physics, init_chains

LJ_25_SIGMA = 2.5 / 2 ** (1 / 6)  # 2.5 sigma in units of r_WCA = 2^(1/6) sigma


def wca_params(repel_d_max, eps, shift, cutoff_factor):
    """populations.py:48-58."""
    s6 = repel_d_max ** 6
    fac = 12.0 * s6 * eps
    return np.array([fac, fac * s6, shift, cutoff_factor * repel_d_max - shift])


def box_from_area_fraction(n_chains, chain_len, a, r, phi, dim=2):
    """Inverse of config.py:80-84: each chain excludes r * (a (L-1) + r)."""
    area = n_chains * r * (a * (chain_len - 1) + r) / phi
    return area ** (1.0 / dim) if dim == 2 else (area * r) ** (1.0 / 3.0)


def grid_constants(rest_length, chain_len, rc, grid_factor=2.0, timestep_safety_factor=0.025,
                   rk_cheat=1.0):
    """config.py:147-193: (max_dist_per_step, n_steps_update_grid_list)."""
    r_int = max(2 * rest_length if chain_len > 1 else 0.0, rc)
    n_upd = max(1, int(1.0 / (timestep_safety_factor * rk_cheat)))
    mdps = 0.5 * (grid_factor - 1) * r_int / n_upd
    return mdps, n_upd


@dataclass
class Workload:
    name: str
    phys: Physics
    pos: np.ndarray
    tangent: np.ndarray
    chain_id: np.ndarray
    rest_lengths: np.ndarray
    v_tan_angle: np.ndarray
    v0_per_chain: np.ndarray
    time_vals: np.ndarray  # for the first call: dt = dt0
    steps: int

    @property
    def n(self):
        return self.phys.n_beads


def chains(name, dim, n_chains, chain_len, *, phi=None, L=None, a=0.5, r=1.0, eps=1.0,
           cutoff_factor=1.0, shift=0.0, k=100.0, kappa=10.0, theta0=0.0, v0=1.0, vtan=0.0,
           T=1e-3, Dr=None, dtmax=1e-3, steps=10000, seed=0, n_nearby=4, tangent_locked=True,
           skin=0.0, min_dist=0.9, min_dist_floor=0.5):
    """A population of equal chains (Equal_Size_Chain_Collection, populations.py:593-611)."""
    if L is None:
        L = box_from_area_fraction(n_chains, chain_len, a, r, phi, dim)
    lens = np.full(n_chains, chain_len, dtype=np.int32)
    pos, tang = init_chains(dim, L, lens, a, min_dist=min_dist, seed=seed,
                            min_dist_floor=min_dist_floor)
    chain_id = np.repeat(np.arange(n_chains, dtype=np.int64), chain_len)
    wca = wca_params(r, eps, shift, cutoff_factor)
    D = T / r
    if Dr is None:
        Dr = 3 * D / (4 * r ** 2)
    get_orient = chain_len > 1 and tangent_locked
    spring = np.array([k, kappa, theta0, 0.0, float(get_orient), 0.1])
    vt = np.broadcast_to(np.asarray(vtan, dtype=np.float64), (n_chains,)).copy()
    phys = Physics(dim=dim, n_beads=n_chains * chain_len, n_chains=n_chains, L=L, D=D, Dr=Dr,
                   radius=r, wca=wca, spring=spring, v0=v0, n_nearby=n_nearby,
                   has_springs=chain_len > 1, has_wca=True, has_propel=True, skin=skin,
                   seed=seed)
    mdps, n_upd = grid_constants(a, chain_len, wca[3])
    tv = np.array([1e-4 * dtmax, dtmax, mdps, 1.0, -1.0, n_upd, n_upd, steps], dtype=np.float64)
    return Workload(name, phys, pos, tang, chain_id, np.full(n_chains, a), vt,
                    np.full(n_chains, v0), tv, steps)


def walled(w, walls, zone, margin=0.8):
    """The workload moved inside WCA walls (forces.py:405-435): the chains that
    do not cross the periodic faces are kept, unchanged, and the box grows by
    zone + margin on each side, so the beads nearest the faces start `margin`
    from a wall."""
    import copy
    L = w.phys.L
    m = w.n // w.phys.n_chains
    pos = w.pos.reshape(w.phys.n_chains, m, -1)
    keep = np.all(np.abs(np.diff(pos, axis=1)) < L / 2, axis=(1, 2))
    keep = np.nonzero(keep)[0]
    pad = zone + margin
    out = copy.deepcopy(w)
    p = out.phys
    p.L = L + 2 * pad
    p.n_chains = len(keep)
    p.n_beads = len(keep) * m
    p.walls = walls
    p.wall_zone = zone
    out.pos = (pos[keep] + pad).reshape(-1, pos.shape[2])
    out.tangent = w.tangent.reshape(w.phys.n_chains, m, -1)[keep].reshape(-1, pos.shape[2])
    out.chain_id = np.repeat(np.arange(len(keep), dtype=np.int64), m)
    out.rest_lengths = w.rest_lengths[keep]
    out.v_tan_angle = w.v_tan_angle[keep]
    out.v0_per_chain = w.v0_per_chain[keep]
    return out


def chains_4x7(seed=0, steps=10000):
    """2D box 20, 4 chains x 7 beads, WCA, k=100, no bending, v0=1, T=0.01, dt=1e-3."""
    return chains("chains_4x7", 2, 4, 7, L=20.0, kappa=0.0, T=0.01, dtmax=1e-3, steps=steps,
                  seed=seed)


def chains_1000x20(seed=0, steps=100000, n_chains=1000):
    """2D, 1000 chains x 20 (20k beads), phi=0.3, WCA, kappa=10, chirality 0.2, T=0.01."""
    return chains("chains_1000x20", 2, n_chains, 20, phi=0.3, kappa=10.0, vtan=0.2, T=0.01,
                  dtmax=1e-3, steps=steps, seed=seed)


def chains_20kx20(seed=0, steps=10000, n_chains=20000):
    """2D, 20k chains x 20 (400k beads), phi=0.4, LJ 2.5 sigma, kappa=5, chirality
    U[-0.5, 0.5] per chain, v0=1, T=0.05, dt=5e-4."""
    rng = np.random.default_rng(seed + 12345)
    vt = rng.uniform(-0.5, 0.5, n_chains)
    return chains("chains_20kx20", 2, n_chains, 20, phi=0.4, cutoff_factor=LJ_25_SIGMA, kappa=5.0,
                  vtan=vt, T=0.05, dtmax=5e-4, steps=steps, seed=seed)


def uneven_replicas(seed=0, n_replicas=256, steps=10000):
    """Ensemble: chain counts log-U[50, 2000], lengths U{5..40}, chirality U[-0.5, 0.5],
    speed U[0.1, 2], area fraction U[0.1, 0.5], WCA, dt=1e-3."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_replicas):
        nc = int(np.exp(rng.uniform(np.log(50), np.log(2000))))
        ln = int(rng.integers(5, 41))
        vt = float(rng.uniform(-0.5, 0.5))
        v0 = float(rng.uniform(0.1, 2.0))
        phi = float(rng.uniform(0.1, 0.5))
        out.append(chains(f"rep{i}", 2, nc, ln, phi=phi, vtan=vt, v0=v0, T=1e-3, dtmax=1e-3,
                          steps=steps, seed=seed * 100003 + i))
    return out


def aligned_rods(w, seed=0):
    """w with an aligned nematic start: every chain a straight rod along +x or
    -x (probability 1/2 each), on a lattice of rod-sized cells filled in random
    order, so that the nematic order is 1 and the polar order near 0 at t = 0.
    Cells are at least one rod length + 1.2 wide and 1.05 high, so no two
    beads of different chains start closer than the WCA range r = 1."""
    rng = np.random.default_rng(seed)
    L, C = w.phys.L, w.phys.n_chains
    m = w.n // C
    a = float(w.rest_lengths[0])
    nx = int(L // ((m - 1) * a + 1.2))
    ny = -(-C // nx)
    dy = L / ny
    if w.phys.dim != 2 or nx < 1 or dy <= 1.05:
        raise ValueError("aligned_rods: 2D box too small for the rods")
    pos = np.zeros((w.n, 2))
    tang = np.zeros((w.n, 2))
    k = np.arange(m)
    for c, cell in enumerate(rng.permutation(nx * ny)[:C]):
        sgn = 1.0 if rng.random() < 0.5 else -1.0
        x0 = (cell % nx) * (L / nx) + 0.1
        sl = slice(c * m, (c + 1) * m)
        pos[sl, 0] = np.mod(x0 + (k if sgn > 0 else m - 1 - k) * a, L)
        pos[sl, 1] = (cell // nx + 0.5) * dy
        tang[sl, 0] = sgn
    w.pos, w.tangent = pos, tang
    return w
