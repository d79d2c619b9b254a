"""Drop-in replacement of updates.SRK2_force_calc (updates.py:132-267).

The reference's driver calls ``updates.SRK2_force_calc(pops, status, sim)``
once per output chunk (control.py:84); that function marshals every array
into ``SRK.SRK_jit`` and writes the step status back. The function built here
has the same signature and the same effects on ``pops`` and ``status``, with
the step loop running on the GPU through libgcr.so:

    import updates, config
    from gcr_amd.dropin import make_SRK2_force_calc
    updates.SRK2_force_calc = make_SRK2_force_calc(config)

Device state stays resident between calls; positions are uploaded again only
when the host changed them. Scope: one moving population of chains (ABPs are
chains of one bead) with the "springs", "WCA", "self-propulsion" and "walls"
forces, SRK_order 2 or 4. Anything else raises NotImplementedError naming the
feature.
"""
import numpy as np

from . import engine as E

SUPPORTED = ("springs", "WCA", "self-propulsion", "walls")


def _force_list(pop, repulsive_only):
    """updates.interpop_forces (updates.py:274-312) for a population with itself."""
    if pop.name in getattr(pop, "special_interactions", {}):
        flist = pop.special_interactions[pop.name]
    else:
        flist = pop.self_interactions()
    out = {}
    for name, params in flist:
        params = np.array(params, dtype=np.float64)
        if name == "WCA" and repulsive_only and params[0] != 0:
            params[3] = (params[1] / params[0]) ** (1 / 6)  # updates.py:304-310
        out[name] = params
    return out


def _mix(seed, k):
    z = (int(seed) * 0x9E3779B97F4A7C15 + k * 0xBF58476D1CE4E5B9 + 1) & ((1 << 64) - 1)
    z ^= z >> 31
    return z


class _Resident:
    """GPU system of one population plus what is needed to detect host edits."""

    def __init__(self):
        self.system = None
        self.sig = None
        self.generation = 0
        self.pos = None
        self.tangent = None


def make_SRK2_force_calc(config, precision=E.GCR_FP64, seed=None):
    """Return a SRK2_force_calc(pops, status, sim) bound to the reference's
    `config` module (its globals Lx, dim, SRK_order, max_dist_per_step,
    RK_cheat_factor, n_steps_update_grid_list, walls are read each call)."""
    resident = {}

    def SRK2_force_calc(pops, status, sim):
        moving = [p for p in pops if not p.fixed]
        if len(pops) != 1 or len(moving) != 1:
            raise NotImplementedError("GPU path: exactly one (moving) population; "
                                      "multi-population interactions are not ported")
        pop = moving[0]
        order = int(getattr(config, "SRK_order", 2))
        if order not in (2, 4):
            raise NotImplementedError(f"GPU path: SRK_order 2 or 4 (not {order})")
        if not hasattr(pop, "chain_id"):
            raise NotImplementedError("GPU path: populations with chain_id (chains / ABPs) only")
        repulsive_only = bool(getattr(pop, "b_init_growing", False))
        forces = _force_list(pop, repulsive_only)
        unknown = [f for f in forces if f not in SUPPORTED]
        if unknown:
            raise NotImplementedError(f"GPU path: forces {unknown} are not ported")
        longest = int(getattr(pop, "longest_chain_length", 1))
        has_springs = "springs" in forces and longest > 1
        has_wca = "WCA" in forces
        has_propel = "self-propulsion" in forces and bool(status.b_calc_active_forces)
        dim = int(config.dim)
        L = float(config.Lx)
        n = int(pop.total_pop)
        n_chains = int(np.max(pop.chain_id)) + 1
        spring = forces["springs"][:6] if has_springs else np.zeros(6)
        spring = np.concatenate([spring, np.zeros(6 - len(spring))])
        wca = forces["WCA"][:4] if has_wca else np.array([0.0, 0.0, 0.0, 1.0])
        act = forces.get("self-propulsion", np.zeros(3))
        v0, vtan = float(act[0]), float(act[2]) if len(act) > 2 else 0.0
        vtan_chain = getattr(pop, "v_tan_angle_per_chain", None)
        if vtan_chain is None:
            vtan_chain = np.full(n_chains, vtan)
        n_nearby = int(getattr(pop, "n_nearby_beads_dont_repel", 0))
        rest = np.asarray(pop.rest_lengths, dtype=np.float64)
        rod = float(getattr(pop, "rigid_rod_projection_factor", 0.0))
        walls, wall_zone, wall_cutoff = 0, 0.0, 0.0
        if "walls" in forces:
            # wall_params = WCA params + config.walls (populations.py:222-223); the
            # walls sit grid_spacing * grid_window / 2 inside the faces, with the
            # grid window of SRK.py:99-103 (forces.py:417)
            wp = forces["walls"]
            walls = sum(1 << i for i in range(int(config.dim)) if wp[4 + i] != 0)
            spacing = float(config.Lx) / int(config.n_grid)
            nbor = float(forces["WCA"][3]) if "WCA" in forces else 0.0
            window = max(1, int(np.ceil(float(config.grid_factor) * nbor / spacing)))
            wall_zone = spacing * window / 2
            wall_cutoff = float(wp[3])
        # cells of at least L/512 (2D) or L/64 (3D) per side, so that the tiny
        # cutoffs of the growth phase (updates.py:304-310) do not ask for a
        # grid of millions of cells
        rc = float(wca[3])
        skin = max(0.3 * rc, L / (512 if dim == 2 else 64) - rc)
        phys = E.Physics(dim=dim, n_beads=n, n_chains=n_chains, L=L, D=float(pop.D),
                         Dr=float(max(pop.Dr, 0.0)), radius=float(pop.repel_d_max), wca=wca,
                         spring=spring, v0=v0, n_nearby=n_nearby, has_springs=has_springs,
                         has_wca=has_wca, has_propel=has_propel, skin=skin, rod=rod,
                         srk_order=order, walls=walls, wall_zone=wall_zone,
                         wall_cutoff=wall_cutoff)
        sig = (n, dim, L, phys.D, phys.Dr, phys.radius, n_nearby, tuple(wca), tuple(spring), v0,
               has_springs, has_wca, has_propel, rod, order, walls, wall_zone, wall_cutoff,
               rest.tobytes(),
               np.asarray(vtan_chain, dtype=np.float64).tobytes(),
               np.asarray(pop.chain_id).tobytes())
        R = resident.setdefault(id(pop), _Resident())
        if R.sig != sig:
            if R.system is not None:
                R.system.close()
            R.generation += 1
            base = sim.get("seed", 0) if seed is None else seed
            phys.seed = _mix(base, R.generation)
            R.system = E.System(phys, pop.pos, pop.tangent_list, pop.chain_id, rest,
                                vtan_chain, precision=precision)
            R.sig = sig
        else:
            if not np.array_equal(pop.pos, R.pos):
                R.system.upload(pos=pop.pos)
            if not np.array_equal(pop.tangent_list, R.tangent):
                R.system.upload(tangent=pop.tangent_list)
        # time_vals exactly as updates.py:138-146
        if sim["save_state_every"] > 0:
            max_dt_usable = status.next_save_time - status.t
        else:
            max_dt_usable = -1
        n_upd = config.n_steps_update_grid_list
        time_vals = np.array([status.dt, status.dtmax, config.max_dist_per_step,
                              config.RK_cheat_factor, max_dt_usable,
                              n_upd - status.update_grid_list_counter, n_upd,
                              status.next_output_step - status.stepnum], dtype=np.float64)
        st = R.system.run(time_vals)
        out = R.system.download()
        pop.pos[:] = out["pos"]
        pop.pos1[:] = out["pos1"]
        pop.tangent_list[:] = out["tangent"]
        pop.forces[:] = out["forces"]
        pop.where_active[:] = -1  # SRK.py:132
        R.pos = pop.pos.copy()
        R.tangent = pop.tangent_list.copy()
        # status write-back, updates.py:249-256
        status.dt, status.dt_used, steps_done, status.max_forces_strength = st[:4]
        status.stepnum += int(steps_done)
        status.t += status.dt_used
        status.update_grid_list_counter = max(
            1, (int(steps_done) + status.update_grid_list_counter) % n_upd)

    SRK2_force_calc.resident = resident
    return SRK2_force_calc
