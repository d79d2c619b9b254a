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
when the host changed them. Scope: populations of chains (ABPs are chains of
one bead) with the "springs", "WCA", "self-propulsion" and "walls" forces,
SRK_order 2 or 4 for one population; several populations (moving or fixed,
interacting through WCA laws: interactions() or special_interactions, as
interpop_forces gives them, updates.py:274-312) with SRK_order 2 and no rigid
rod. Anything else raises NotImplementedError naming the feature.
"""
import numpy as np

from . import engine as E

SUPPORTED = ("springs", "WCA", "self-propulsion", "walls", "active_sliding_and_WCA")


def _force_list(pop, repulsive_only, other=None):
    """updates.interpop_forces (updates.py:274-312) for a population with
    itself, or with `other`."""
    other = pop if other is None else other
    if other.name in getattr(pop, "special_interactions", {}):
        flist = pop.special_interactions[other.name]
    elif other is pop:
        flist = pop.self_interactions()
    else:
        flist = pop.interactions(other)
    out = {}
    for name, params in flist:
        params = np.array(params, dtype=np.float64)
        if name in ("WCA", "active_sliding_and_WCA") and repulsive_only and params[0] != 0:
            params[3] = (params[1] / params[0]) ** (1 / 6)  # updates.py:304-310
        out[name] = params
    return out


def chain_chunks(n_threads, chain_id):
    """The thread chunk of every chain when updates.chunk_chain
    (updates.py:323-344) splits a population among n_threads: a cut about
    every ceil(n / n_threads) beads, moved on to the next chain boundary."""
    chain_id = np.asarray(chain_id)
    n = len(chain_id)
    chunklen = int(np.ceil(n / n_threads))
    splits, first, sii = [], 0, 0
    while first < n and sii < n_threads - 1:
        if first > 0:
            splits.append(first)
            sii += 1
        test = min(max(first, (sii + 1) * chunklen - 1), n - 1)
        last = chain_id[test]
        while test < n and chain_id[test] == last:
            test += 1
        first = test
    bead_chunk = np.zeros(n, dtype=np.int32)
    for s_ in splits:
        bead_chunk[s_:] += 1
    n_chains = int(chain_id.max()) + 1 if n else 0
    out = np.zeros(n_chains, dtype=np.int32)
    out[chain_id[::-1]] = bead_chunk[::-1]  # the chunk of each chain's first bead
    return out


def _thread_counts(config, sim, n_pairs):
    """updates.py:172-182: threads of the first population pair and of the others."""
    n = int(getattr(config, "n_threads", 1))
    if n < n_pairs:
        n = n_pairs
    n = min(n, max(n_pairs, int(sim.get("n_threads", 1))))
    config.n_threads = n
    per = max(1, n // n_pairs)
    return n - (n_pairs - 1) * per, per


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


def _time_vals(config, status, sim):
    """time_vals exactly as updates.py:138-146."""
    if sim["save_state_every"] > 0:
        max_dt_usable = status.next_save_time - status.t
    else:
        max_dt_usable = -1
    n_upd = config.n_steps_update_grid_list
    return np.array([status.dt, status.dtmax, config.max_dist_per_step, config.RK_cheat_factor,
                     max_dt_usable, n_upd - status.update_grid_list_counter, n_upd,
                     status.next_output_step - status.stepnum], dtype=np.float64)


def _write_status(config, status, st):
    """Status write-back, updates.py:249-256."""
    n_upd = config.n_steps_update_grid_list
    status.dt, status.dt_used, steps_done, status.max_forces_strength = st[:4]
    status.stepnum += int(steps_done)
    status.t += status.dt_used
    status.update_grid_list_counter = max(
        1, (int(steps_done) + status.update_grid_list_counter) % n_upd)


def _motor_population(pop, config, fixed):
    """A Motor_Population (populations.py:617-660): single beads with no
    forces of their own, diffusing ("diffuse", SRK.py:389-396) unless fixed."""
    dim = int(config.dim)
    L = float(config.Lx)
    n = int(pop.total_pop)
    diffuse = dict(pop.self_interactions()).get("diffuse")
    still = bool(diffuse[0]) if diffuse is not None else True
    phys = E.Physics(dim=dim, n_beads=n, n_chains=n, L=L, D=0.0 if fixed else float(pop.D),
                     Dr=0.0, radius=1.0, wca=np.array([0.0, 0.0, 0.0, 1.0]), spring=np.zeros(6),
                     v0=0.0, n_nearby=-1, has_springs=False, has_wca=False, has_propel=False,
                     skin=max(0.3, L / (512 if dim == 2 else 64) - 1.0), motor=True,
                     motor_still=still, motor_max_dist=float(pop.motor_max_dist),
                     motor_imparts=float(pop.imparts_force), k_on=float(pop.k_on),
                     k_off=float(pop.k_off))
    chain_id = np.arange(n, dtype=np.int64)
    sig = ("motor", n, dim, L, phys.D, still, phys.motor_max_dist, phys.motor_imparts,
           phys.k_on, phys.k_off, fixed)
    return phys, np.zeros(n), np.ones(n), chain_id, sig


def _population(pop, config, status, order, fixed=False):
    """Physics, per-chain chirality, rest lengths, chain ids and the
    signature of one population, from its force list (SRK.py:46-97)."""
    if hasattr(pop, "k_on") and hasattr(pop, "motor_max_dist"):
        return _motor_population(pop, config, fixed)
    dim = int(config.dim)
    L = float(config.Lx)
    n = int(pop.total_pop)
    chain_id = np.asarray(getattr(pop, "chain_id", np.arange(n)), dtype=np.int64)
    n_chains = int(np.max(chain_id)) + 1 if n else 0
    radius = float(getattr(pop, "repel_d_max", 1.0))
    if fixed:
        # a fixed population only pushes the others; its cells follow its own
        # WCA cutoff when it has one
        wca = np.asarray(getattr(pop, "WCA_params", [0.0, 0.0, 0.0, 1.0]), dtype=np.float64)[:4]
        phys = E.Physics(dim=dim, n_beads=n, n_chains=n_chains, L=L, D=0.0, Dr=0.0,
                         radius=radius, wca=wca, spring=np.zeros(6), v0=0.0, n_nearby=-1,
                         has_springs=False, has_wca=True, has_propel=False,
                         skin=max(0.3 * float(wca[3]), L / (512 if dim == 2 else 64) - wca[3]))
        rest = np.ones(n_chains)
        vt = np.zeros(n_chains)
        sig = ("fixed", n, dim, L, tuple(wca), chain_id.tobytes())
        return phys, vt, rest, chain_id, sig
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
    spring = forces["springs"][:6] if has_springs else np.zeros(6)
    spring = np.concatenate([spring, np.zeros(6 - len(spring))])
    wca = forces["WCA"][:4] if has_wca else np.array([0.0, 0.0, 0.0, 1.0])
    act = forces.get("self-propulsion", np.zeros(3))
    v0, vtan = float(act[0]), float(act[2]) if len(act) > 2 else 0.0
    sliding_maxdist = sliding_v0 = 0.0
    if "active_sliding_and_WCA" in forces:
        # WCA params + active_params (populations.py:580-591): the pair law,
        # and with active forces on, the push v0 per antiparallel partner
        # closer than sliding_partner_maxdist (SRK.py:73-80)
        sp = forces["active_sliding_and_WCA"]
        wca = np.asarray(sp[:4], dtype=np.float64)
        has_wca = True
        vtan = float(sp[6]) if len(sp) > 6 else 0.0
        if bool(status.b_calc_active_forces):
            sliding_v0, sliding_maxdist = float(sp[4]), float(sp[5])
            if order == 4:
                raise NotImplementedError("GPU path: active sliding with SRK_order 4")
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
        walls = sum(1 << i for i in range(dim) if wp[4 + i] != 0)
        spacing = L / int(config.n_grid)
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
                     Dr=float(max(pop.Dr, 0.0)), radius=radius, wca=wca,
                     spring=spring, v0=v0, n_nearby=n_nearby, has_springs=has_springs,
                     has_wca=has_wca, has_propel=has_propel, skin=skin, rod=rod,
                     srk_order=order, walls=walls, wall_zone=wall_zone,
                     wall_cutoff=wall_cutoff, sliding_maxdist=sliding_maxdist,
                     sliding_v0=sliding_v0)
    sig = (n, dim, L, phys.D, phys.Dr, phys.radius, n_nearby, tuple(wca), tuple(spring), v0,
           has_springs, has_wca, has_propel, rod, order, walls, wall_zone, wall_cutoff,
           sliding_maxdist, sliding_v0,
           rest.tobytes(), np.asarray(vtan_chain, dtype=np.float64).tobytes(),
           chain_id.tobytes())
    return phys, vtan_chain, rest, chain_id, sig


def _cross_laws(pops, moving):
    """interpop_forces(P, Q) for every moving P and every other Q: the WCA law
    of Q acting on P, or "motor" when Q's motors push P (updates.py:274-312)."""
    laws = {}
    for a, P in enumerate(pops):
        if not moving[a] or not hasattr(P, "forces"):
            continue  # SRK2_force_calc gives populations without forces only themselves
        growing = bool(getattr(P, "b_init_growing", False))
        for b, Q in enumerate(pops):
            if b == a:
                continue
            fl = _force_list(P, growing or bool(getattr(Q, "b_init_growing", False)), Q)
            unknown = [f for f in fl if f not in ("WCA", "motor_push")]
            if unknown:
                raise NotImplementedError(
                    f"GPU path: forces {unknown} between populations are not ported")
            if "WCA" in fl and "motor_push" in fl:
                raise NotImplementedError("GPU path: WCA and motor_push between one pair")
            if "WCA" in fl:
                laws[(a, b)] = np.asarray(fl["WCA"][:4], dtype=np.float64)
            if "motor_push" in fl:
                laws[(a, b)] = "motor"
    return laws


class _Group:
    """Device pool holding several populations as one group."""

    def __init__(self):
        self.pool = None
        self.rids = None
        self.sig = None
        self.generation = 0
        self.pos = None
        self.tangent = None


def make_SRK2_force_calc(config, precision=E.GCR_FP64, seed=None):
    """Return a SRK2_force_calc(pops, status, sim) bound to the reference's
    `config` module (its globals Lx, dim, SRK_order, max_dist_per_step,
    RK_cheat_factor, n_steps_update_grid_list, walls, n_grid, grid_factor are
    read each call)."""
    resident = {}

    def one_population(pop, status, sim, order):
        phys, vtan_chain, rest, chain_id, sig = _population(pop, config, status, order)
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
        if phys.sliding_maxdist > 0:
            first, _ = _thread_counts(config, sim, 1)
            R.system.set_chunks(chain_chunks(first, chain_id))
        else:
            if not np.array_equal(pop.pos, R.pos):
                R.system.upload(pos=pop.pos)
            if not np.array_equal(pop.tangent_list, R.tangent):
                R.system.upload(tangent=pop.tangent_list)
        st = R.system.run(_time_vals(config, status, sim))
        out = R.system.download()
        pop.pos[:] = out["pos"]
        pop.pos1[:] = out["pos1"]
        pop.tangent_list[:] = out["tangent"]
        pop.forces[:] = out["forces"]
        # SRK.py:132, forces.py:119: -1, plus the sliding partners of the last step
        pop.where_active[:] = R.system.where_active() if phys.sliding_maxdist > 0 else -1
        R.pos = pop.pos.copy()
        R.tangent = pop.tangent_list.copy()
        _write_status(config, status, st)

    def several_populations(pops, status, sim, order):
        if order != 2:
            raise NotImplementedError("GPU path: several populations with SRK_order 2 only")
        moving = [not p.fixed for p in pops]
        parts = [_population(p, config, status, order, fixed=not moving[k])
                 for k, p in enumerate(pops)]
        if any(ph.rod != 0.0 and moving[k] for k, (ph, *_rest) in enumerate(parts)):
            raise NotImplementedError("GPU path: the rigid rod with several populations")
        laws = _cross_laws(pops, moving)
        sig = (tuple(pt[4] for pt in parts), tuple(moving),
               tuple(sorted((k, v if isinstance(v, str) else tuple(v)) for k, v in laws.items())))
        key = tuple(id(p) for p in pops)
        G = resident.setdefault(key, _Group())
        tangents = [getattr(p, "tangent_list", None) for p in pops]
        if G.sig != sig:
            if G.pool is not None:
                G.pool.close()
            G.generation += 1
            base = sim.get("seed", 0) if seed is None else seed
            dim = int(config.dim)
            nb = sum(int(p.total_pop) for p in pops)
            nch = sum(pt[0].n_chains for pt in parts)
            cells = sum(E.replica_cells(pt[0]) for pt in parts)
            G.pool = E.Pool(dim, precision, bead_capacity=nb, cell_capacity=cells + len(pops),
                            chain_capacity=nch, replica_capacity=len(pops))
            G.rids = []
            for k, (p, (phys, vt, rest, cid, _s)) in enumerate(zip(pops, parts)):
                phys.seed = _mix(base, G.generation * 64 + k)
                t = tangents[k] if tangents[k] is not None else np.tile(
                    np.eye(dim)[0], (int(p.total_pop), 1))
                G.rids.append(G.pool.add(phys, p.pos, t, cid, rest, vt))
            G.pool.group(G.rids, laws, [k for k in range(len(pops)) if not moving[k]])
            G.sig = sig
            # thread chunks for active sliding: updates.py:172-200 gives the first
            # moving population's own pair the leftover threads
            n_pairs = sum(len(pops) for k in range(len(pops)) if moving[k])
            first, per = _thread_counts(config, sim, n_pairs)
            first_moving = moving.index(True)
            for k, (phys, _v, _r, cid, _s) in enumerate(parts):
                if moving[k] and phys.sliding_maxdist > 0:
                    G.pool.set_chunks(G.rids[k],
                                      chain_chunks(first if k == first_moving else per, cid))
        else:
            for k, p in enumerate(pops):
                if not np.array_equal(p.pos, G.pos[k]):
                    G.pool.upload(G.rids[k], pos=p.pos)
                if tangents[k] is not None and not np.array_equal(tangents[k], G.tangent[k]):
                    G.pool.upload(G.rids[k], tangent=tangents[k])
        tv = _time_vals(config, status, sim)
        for k in range(len(pops)):
            if moving[k]:
                G.pool.begin(G.rids[k], tv)
        G.pool.run()
        st = None
        for k, p in enumerate(pops):
            if not moving[k]:
                if parts[k][0].motor:  # a fixed motor population's bindings persist
                    p.where_active[:] = G.pool.where_active(G.rids[k])
                continue
            out = G.pool.download(G.rids[k])
            p.pos[:] = out["pos"]
            p.pos1[:] = out["pos1"]
            if tangents[k] is not None:
                p.tangent_list[:] = out["tangent"]
            if hasattr(p, "forces"):
                p.forces[:] = out["forces"]
            # SRK.py:132, forces.py:119, 262-281: -1 plus sliding partners and
            # motor bindings; for motors, the bead each one pushes
            p.where_active[:] = G.pool.where_active(G.rids[k])
            if st is None:
                s_ = G.pool.status(G.rids[k])
                st = np.array([s_["dt"], s_["elapsed"], s_["steps"], s_["max_forces_strength"]])
        G.pos = [p.pos.copy() for p in pops]
        G.tangent = [None if t is None else t.copy() for t in tangents]
        _write_status(config, status, st)

    def SRK2_force_calc(pops, status, sim):
        moving = [p for p in pops if not p.fixed]
        if not moving:
            raise NotImplementedError("GPU path: no moving population")
        order = int(getattr(config, "SRK_order", 2))
        if order not in (2, 4):
            raise NotImplementedError(f"GPU path: SRK_order 2 or 4 (not {order})")
        if len(pops) == 1:
            one_population(pops[0], status, sim, order)
        else:
            several_populations(pops, status, sim, order)

    SRK2_force_calc.resident = resident
    return SRK2_force_calc
