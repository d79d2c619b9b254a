"""Python handles over libgcr.so: a pool of replicas and the single-system form.

Device state stays resident between calls; host arrays cross only at
add/upload/download. Every call goes to the HIP library (no CPU path).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import GCR_FP64, GCR_MIXED, ReplicaDesc, check, f64, ptr

__all__ = ["Physics", "Pool", "System", "GCR_FP64", "GCR_MIXED", "init_chains"]


@dataclass
class Physics:
    """One population's physics as SRK_jit receives it (gcr_replica_desc)."""

    dim: int
    n_beads: int
    n_chains: int
    L: float
    D: float
    Dr: float
    radius: float = 1.0
    wca: np.ndarray = field(default_factory=lambda: np.array([12.0, 12.0, 0.0, 1.0]))
    spring: np.ndarray = field(default_factory=lambda: np.array([100.0, 10.0, 0.0, 0.0, 1.0, 0.1]))
    v0: float = 1.0
    n_nearby: int = 4
    has_springs: bool = True
    has_wca: bool = True
    has_propel: bool = True
    skin: float = 0.0
    seed: int = 0
    rod: float = 0.0  # rigid_rod_projection_factor
    srk_order: int = 2  # config.SRK_order (2 or 4)
    walls: int = 0  # config.walls as bits: x 1, y 2, z 4
    wall_zone: float = 0.0  # forbidden_zone_width (forces.py:417)
    wall_cutoff: float = 0.0  # wall_params[3]; <= 0: wca[3]
    sliding_maxdist: float = 0.0  # active sliding partner range (0: off)
    sliding_v0: float = 0.0  # active sliding push per partner
    motor: bool = False  # a motor population
    motor_still: bool = True  # b_motors_dont_diffuse_while_pushing
    motor_max_dist: float = 0.0
    motor_imparts: float = 0.0  # motor_imparts_force
    k_on: float = 0.0
    k_off: float = 0.0
    motor_reach: float = 0.0  # <= 0: motor_max_dist

    def desc(self):
        d = ReplicaDesc()
        d.n_beads = int(self.n_beads)
        d.n_chains = int(self.n_chains)
        d.n_nearby = int(self.n_nearby)
        d.has_springs = int(bool(self.has_springs))
        d.has_wca = int(bool(self.has_wca))
        d.has_propel = int(bool(self.has_propel))
        d.L = float(self.L)
        d.D = float(self.D)
        d.Dr = float(self.Dr)
        d.radius = float(self.radius)
        for i in range(4):
            d.wca[i] = float(self.wca[i])
        for i in range(6):
            d.spring[i] = float(self.spring[i])
        d.v0 = float(self.v0)
        d.skin = float(self.skin)
        d.seed = int(self.seed) & ((1 << 64) - 1)
        d.rod = float(self.rod)
        d.srk_order = int(self.srk_order)
        d.walls = int(self.walls)
        d.wall_zone = float(self.wall_zone)
        d.wall_cutoff = float(self.wall_cutoff)
        d.sliding_maxdist = float(self.sliding_maxdist)
        d.sliding_v0 = float(self.sliding_v0)
        d.motor = int(bool(self.motor))
        d.motor_still = int(bool(self.motor_still))
        d.motor_max_dist = float(self.motor_max_dist)
        d.motor_imparts = float(self.motor_imparts)
        d.k_on = float(self.k_on)
        d.k_off = float(self.k_off)
        d.motor_reach = float(self.motor_reach)
        return d


def _time_vals(tv):
    tv = f64(tv)
    assert tv.shape == (8,)
    return tv


class Pool:
    """Many replicas of one dimensionality advanced in the same launches."""

    def __init__(self, dim, precision=GCR_FP64, bead_capacity=1 << 20, cell_capacity=1 << 20,
                 chain_capacity=1 << 18, replica_capacity=64):
        self._L = _lib.load()
        self.dim = dim
        self.precision = precision
        h = C.c_void_p()
        check(self._L.gcr_pool_create(dim, precision, int(bead_capacity), int(cell_capacity),
                                      int(chain_capacity), int(replica_capacity), C.byref(h)))
        self.h = h
        self.n = {}

    def close(self):
        if self.h:
            check(self._L.gcr_pool_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, steps_per_batch, rebuild_every):
        check(self._L.gcr_pool_set_schedule(self.h, int(steps_per_batch), int(rebuild_every)))

    def set_tile(self, side):
        """Tile edge of the LDS force kernels (0: automatic), or for 2D a
        (cells in x, cells in y) rectangle."""
        if isinstance(side, (tuple, list)):
            check(self._L.gcr_pool_set_tile_shape(self.h, int(side[0]), int(side[1])))
        else:
            check(self._L.gcr_pool_set_tile(self.h, int(side)))

    def set_tile_threads(self, threads):
        """Threads per tile workgroup of the 2D force kernels (0, 256, 512 or 1024)."""
        check(self._L.gcr_pool_set_tile_threads(self.h, int(threads)))

    def add(self, phys, pos, tangent, chain_id, rest_lengths, v_tan_angle=None,
            v0_per_chain=None):
        pos = f64(pos)
        tangent = f64(tangent)
        cid = np.ascontiguousarray(chain_id, dtype=np.int64)
        rl = f64(rest_lengths)
        vt = f64(v_tan_angle) if v_tan_angle is not None else np.zeros(phys.n_chains)
        v0c = f64(v0_per_chain)
        assert pos.shape == (phys.n_beads, self.dim)
        d = phys.desc()
        rid = C.c_int32()
        check(self._L.gcr_replica_add(self.h, C.byref(d), ptr(pos), ptr(tangent), ptr(cid),
                                      ptr(rl), ptr(vt), ptr(v0c), C.byref(rid)))
        self.n[rid.value] = phys.n_beads
        return rid.value

    def remove(self, rid):
        check(self._L.gcr_replica_remove(self.h, int(rid)))
        self.n.pop(rid, None)

    def set_physics(self, rid, phys, rest_lengths=None, v0_per_chain=None):
        """Change a replica's physics in place between calls (same beads,
        chains and box; the cutoff must fit the cell grid made at add)."""
        d = phys.desc()
        rl, v0c = f64(rest_lengths), f64(v0_per_chain)
        check(self._L.gcr_replica_set_physics(self.h, int(rid), C.byref(d), ptr(rl), ptr(v0c)))

    def upload(self, rid, pos=None, tangent=None):
        pos, tangent = f64(pos), f64(tangent)
        check(self._L.gcr_replica_upload(self.h, int(rid), ptr(pos), ptr(tangent)))

    def begin(self, rid, time_vals, xi=None):
        tv = _time_vals(time_vals)
        xi = f64(xi)
        steps = 0 if xi is None else xi.shape[0]
        if xi is not None:
            assert xi.shape[1:] == (self.n[rid], 2 * self.dim)
        check(self._L.gcr_replica_begin(self.h, int(rid), ptr(tv), ptr(xi), int(steps)))

    def run(self, max_batches=0):
        nr = C.c_int32()
        check(self._L.gcr_pool_run(self.h, int(max_batches), C.byref(nr)))
        return nr.value

    def status(self, rid):
        out = np.zeros(8)
        check(self._L.gcr_replica_status(self.h, int(rid),
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(dt=out[0], elapsed=out[1], steps=int(out[2]), max_forces_strength=out[3],
                    running=bool(out[4]), replays=int(out[5]), step_global=int(out[6]),
                    failed=bool(out[7]))

    def download(self, rid, pos=True, pos1=True, tangent=True, forces=True):
        n = self.n[rid]
        shape = (n, self.dim)
        out = {k: (np.empty(shape) if want else None)
               for k, want in (("pos", pos), ("pos1", pos1), ("tangent", tangent),
                               ("forces", forces))}
        check(self._L.gcr_replica_download(self.h, int(rid), ptr(out["pos"]), ptr(out["pos1"]),
                                           ptr(out["tangent"]), ptr(out["forces"])))
        return {k: v for k, v in out.items() if v is not None}

    def eval_forces(self, rid):
        n = self.n[rid]
        F = np.empty((n, self.dim))
        T = np.empty((n, self.dim))
        check(self._L.gcr_replica_eval_forces(self.h, int(rid), ptr(F), ptr(T)))
        return F, T

    def order_parameters(self, rid):
        """Polar order, nematic order and director angle of a replica's chains,
        computed on the device (no position download)."""
        out = np.zeros(4)
        check(self._L.gcr_replica_order(self.h, int(rid),
                                        out.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(polar=out[0], nematic=out[1], director=out[2], n_chains=int(out[3]))

    def group(self, rids, cross=None, fixed=None):
        """Make replicas `rids` (in population order) the populations of one
        system: cross[(a, b)] = the WCA law 4-vector of population b acting on
        a (interpop_forces), or "motor" when b is a motor population pushing
        a's beads; fixed = the indices of fixed populations."""
        n = len(rids)
        r = np.ascontiguousarray(rids, dtype=np.int32)
        c = np.zeros((n, n, 5))
        for (a, b), law in (cross or {}).items():
            if isinstance(law, str) and law == "motor":
                c[a, b, 0] = 2.0
                continue
            c[a, b, 0] = 1.0
            c[a, b, 1:] = law
        fx = np.zeros(n, dtype=np.int32)
        for k in fixed or ():
            fx[k] = 1
        check(self._L.gcr_pool_group(self.h, ptr(r), n, ptr(c), ptr(fx)))

    def set_chunks(self, rid, chain_chunk):
        """The reference's thread chunk of each chain (active sliding)."""
        c = np.ascontiguousarray(chain_chunk, dtype=np.int32)
        check(self._L.gcr_replica_set_chunks(self.h, int(rid), ptr(c)))

    def where_active(self, rid):
        out = np.empty(self.n[rid], dtype=np.int32)
        check(self._L.gcr_replica_where_active(self.h, int(rid), ptr(out)))
        return out

    def sample(self, rid, reset=False):
        """Appends one row of the replica's observables series on the device
        (gcr_replica_sample): time, polar order, nematic order, director
        angle, chain-centre MSD since the origin, steps."""
        check(self._L.gcr_replica_sample(self.h, int(rid), 1 if reset else 0))

    def series(self, rid):
        """The replica's observables series as a dict of arrays."""
        n = C.c_int32()
        check(self._L.gcr_replica_series(self.h, int(rid), None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 6))
        check(self._L.gcr_replica_series(self.h, int(rid), ptr(out), n.value, C.byref(n)))
        out = out[: n.value]
        return dict(t=out[:, 0], polar=out[:, 1], nematic=out[:, 2], director=out[:, 3],
                    msd=out[:, 4], step=out[:, 5].astype(np.int64))

    def graph_stats(self):
        """Batch-graph work so far: captures, instantiations, in-place
        updates and the host seconds they took (gcr_pool_graph_stats)."""
        out = np.zeros(4)
        check(self._L.gcr_pool_graph_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(captures=int(out[0]), instantiations=int(out[1]), updates=int(out[2]),
                    seconds=float(out[3]))

    def cell_list(self, rid):
        """(cells per side, first slot of each cell, slot of each bead) of
        the replica's last re-sort (gcr_replica_cell_list)."""
        nc = C.c_int32()
        check(self._L.gcr_replica_cell_list(self.h, int(rid), C.byref(nc), None, None))
        ncell = nc.value ** self.dim
        cstart = np.empty(ncell + 1, dtype=np.int32)
        slot = np.empty(self.n[rid], dtype=np.int32)
        check(self._L.gcr_replica_cell_list(self.h, int(rid), C.byref(nc), ptr(cstart), ptr(slot)))
        return nc.value, cstart, slot

    def time_kernels(self, steps):
        out = np.zeros(4)
        check(self._L.gcr_pool_time_kernels(self.h, int(steps),
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(stage0=out[0], stage1=out[1], update=out[2], rebuild=out[3])


class System:
    """Single-system form: the SRK_jit replacement (gcr_create / gcr_run ...)."""

    def __init__(self, phys, pos, tangent, chain_id, rest_lengths, v_tan_angle=None,
                 precision=GCR_FP64):
        self._L = _lib.load()
        self.dim = phys.dim
        self.n = phys.n_beads
        pos, tangent = f64(pos), f64(tangent)
        cid = np.ascontiguousarray(chain_id, dtype=np.int64)
        rl = f64(rest_lengths)
        vt = f64(v_tan_angle) if v_tan_angle is not None else np.zeros(phys.n_chains)
        d = phys.desc()
        h = C.c_void_p()
        check(self._L.gcr_create(C.byref(d), phys.dim, precision, ptr(pos), ptr(tangent),
                                 ptr(cid), ptr(rl), ptr(vt), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(self._L.gcr_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pos=None, tangent=None):
        pos, tangent = f64(pos), f64(tangent)
        check(self._L.gcr_upload(self.h, ptr(pos), ptr(tangent)))

    def run(self, time_vals, xi=None):
        tv = _time_vals(time_vals)
        xi = f64(xi)
        st = np.zeros(6)
        steps = 0 if xi is None else xi.shape[0]
        check(self._L.gcr_run(self.h, ptr(tv), ptr(st), ptr(xi), int(steps)))
        return st

    def download(self):
        out = {k: np.empty((self.n, self.dim)) for k in ("pos", "pos1", "tangent", "forces")}
        check(self._L.gcr_download(self.h, ptr(out["pos"]), ptr(out["pos1"]),
                                   ptr(out["tangent"]), ptr(out["forces"])))
        return out

    def eval_forces(self):
        F = np.empty((self.n, self.dim))
        T = np.empty((self.n, self.dim))
        check(self._L.gcr_eval_forces(self.h, ptr(F), ptr(T)))
        return F, T

    def set_chunks(self, chain_chunk):
        c = np.ascontiguousarray(chain_chunk, dtype=np.int32)
        check(self._L.gcr_replica_set_chunks(self.h, 0, ptr(c)))

    def where_active(self):
        out = np.empty(self.n, dtype=np.int32)
        check(self._L.gcr_replica_where_active(self.h, 0, ptr(out)))
        return out

    def order_parameters(self):
        """Polar order, nematic order, director angle (on the device); the
        system is replica 0 of its one-replica pool."""
        out = np.zeros(4)
        check(self._L.gcr_replica_order(self.h, 0, out.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(polar=out[0], nematic=out[1], director=out[2], n_chains=int(out[3]))


def replica_cells(phys):
    """Cells (plus the end offset) a replica of this physics takes in a pool
    (gcr_replica_cells)."""
    d = phys.desc()
    return int(_lib.load().gcr_replica_cells(int(phys.dim), C.byref(d))) + 1


def init_chains(dim, L, chain_len, bond, min_dist=0.9, seed=0, max_tries=20000,
                min_dist_floor=None, return_min_dist=False):
    """Random straight chains without overlap (gcr_init_chains). With
    min_dist_floor below min_dist, dense boxes that random sequential
    placement cannot fill are completed at a smaller exclusion distance."""
    lib = _lib.load()
    cl = np.ascontiguousarray(chain_len, dtype=np.int32)
    n = int(cl.sum())
    pos = np.empty((n, dim))
    tang = np.empty((n, dim))
    used = C.c_double()
    f = lib.gcr_init_chains
    f.restype = C.c_int
    f.argtypes = [C.c_int32, C.c_double, C.c_int64, C.c_void_p, C.c_double, C.c_double,
                  C.c_double, C.c_uint64, C.c_int64, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_double)]
    floor = min_dist if min_dist_floor is None else min_dist_floor
    rc = f(dim, float(L), len(cl), ptr(cl), float(bond), float(min_dist), float(floor),
           int(seed) & ((1 << 64) - 1), int(max_tries), ptr(pos), ptr(tang), C.byref(used))
    if rc != 0:
        raise _lib.GcrError(
            f"gcr_init_chains failed ({rc}): could not place every chain without overlap")
    return (pos, tang, used.value) if return_min_dist else (pos, tang)
