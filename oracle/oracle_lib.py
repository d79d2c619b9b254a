"""ctypes binding of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg use it as the checker. The product path never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


class OrcSys(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n", C.c_int32),
        ("n_nearby", C.c_int32),
        ("has_springs", C.c_int32),
        ("has_wca", C.c_int32),
        ("has_propel", C.c_int32),
        ("L", C.c_double),
        ("D", C.c_double),
        ("Dr", C.c_double),
        ("radius", C.c_double),
        ("wca", C.c_double * 4),
        ("spring", C.c_double * 6),
        ("v0", C.c_double),
        ("rod", C.c_double),
        ("order", C.c_int32),
        ("walls", C.c_int32),
        ("wall_zone", C.c_double),
        ("wall_cutoff", C.c_double),
        ("sliding_maxdist", C.c_double),
        ("sliding_v0", C.c_double),
        ("chunk", C.c_void_p),
        ("where_active", C.c_void_p),
    ]


class OrcPop(C.Structure):
    _fields_ = [
        ("sys", OrcSys),
        ("fixed", C.c_int32),
        ("pad", C.c_int32),
        ("pos", C.c_void_p),
        ("pos1", C.c_void_p),
        ("tangent", C.c_void_p),
        ("forces", C.c_void_p),
        ("chain_id", C.c_void_p),
        ("rest_lengths", C.c_void_p),
        ("vtan", C.c_void_p),
        ("xi", C.c_void_p),
        ("seed", C.c_uint64),
        ("motor", C.c_int32),
        ("motor_still", C.c_int32),
        ("motor_max_dist_sq", C.c_double),
        ("motor_imparts", C.c_double),
        ("k_on", C.c_double),
        ("k_off", C.c_double),
        ("motor_reach", C.c_double),
        ("motor_bead", C.c_void_p),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        L.orc_pymod.restype = C.c_double
        L.orc_pymod.argtypes = [C.c_double, C.c_double]
        L.orc_wca_coeff.restype = C.c_double
        L.orc_wca_coeff.argtypes = [C.c_double, dp]
        L.orc_wca_params.argtypes = [C.c_double] * 4 + [dp]
        L.orc_chain_forces.argtypes = [C.c_int, C.c_int, dp, ip, C.c_double, dp, dp, dp, dp]
        L.orc_pair_forces_brute.argtypes = [C.c_int, C.c_int, dp, ip, C.c_double, dp,
                                            C.c_int, dp]
        L.orc_self_propel.argtypes = [C.c_int, C.c_int, ip, C.c_double, dp, dp, dp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_bead_noise.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_int64, dp]
        L.orc_srk2_run.restype = C.c_int64
        L.orc_srk2_run.argtypes = [C.POINTER(OrcSys), dp, dp, dp, dp, ip, dp, dp, dp, dp,
                                   C.c_void_p, C.c_uint64, C.c_int64, C.c_int, C.c_int,
                                   C.c_double]
        L.orc_cross_pair_forces.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                            C.c_double, C.c_void_p, C.c_void_p]
        L.orc_multi_srk2_run.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_double]
        L.orc_multi_srk2_run.restype = C.c_int64
        L.orc_motor_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.c_int,
                                        C.c_int]
        L.orc_motor_uniform.restype = C.c_double
        L.orc_walls.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, C.c_int, C.c_double, dp]
        L.orc_force_eval.argtypes = [C.POINTER(OrcSys), dp, ip, dp, dp, dp, dp, C.c_int,
                                     C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def make_sys(sysd):
    s = OrcSys()
    s.dim = sysd["dim"]
    s.n = sysd["n"]
    s.n_nearby = sysd["n_nearby"]
    s.has_springs = int(sysd["has_springs"])
    s.has_wca = int(sysd["has_wca"])
    s.has_propel = int(sysd["has_propel"])
    s.L = sysd["L"]
    s.D = sysd["D"]
    s.Dr = sysd["Dr"]
    s.radius = sysd["radius"]
    for i in range(4):
        s.wca[i] = sysd["wca"][i]
    for i in range(6):
        s.spring[i] = sysd["spring"][i]
    s.v0 = sysd["v0"]
    s.rod = sysd.get("rod", 0.0)
    s.order = sysd.get("order", 2)
    s.walls = sysd.get("walls", 0)
    s.wall_zone = sysd.get("wall_zone", 0.0)
    s.wall_cutoff = sysd.get("wall_cutoff", 0.0)
    s.sliding_maxdist = sysd.get("sliding_maxdist", 0.0)
    s.sliding_v0 = sysd.get("sliding_v0", 0.0)
    ch = sysd.get("chunk")
    s.chunk = None if ch is None else ch.ctypes.data
    wa = sysd.get("where_active")
    s.where_active = None if wa is None else wa.ctypes.data
    return s


def pymod(a, b):
    return lib().orc_pymod(a, b)


def wca_params(r, eps, shift, cf):
    out = np.zeros(4)
    lib().orc_wca_params(r, eps, shift, cf, out)
    return out


def wca_coeff(ds, p):
    return lib().orc_wca_coeff(ds, np.ascontiguousarray(p, dtype=np.float64))


def chain_forces(pos, chain_id, L, sp, rest_lengths, tangent=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, dim = pos.shape
    F = np.zeros_like(pos)
    T = np.full_like(pos, np.nan) if tangent is None else np.array(tangent, dtype=np.float64)
    lib().orc_chain_forces(dim, n, pos, np.ascontiguousarray(chain_id, dtype=np.int64), L,
                           np.ascontiguousarray(sp, dtype=np.float64),
                           np.ascontiguousarray(rest_lengths, dtype=np.float64), T, F)
    return F, T


def pair_forces(pos, chain_id, L, p, n_nearby):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, dim = pos.shape
    F = np.zeros_like(pos)
    lib().orc_pair_forces_brute(dim, n, pos, np.ascontiguousarray(chain_id, dtype=np.int64), L,
                                np.ascontiguousarray(p, dtype=np.float64), n_nearby, F)
    return F


def self_propel(tangent, chain_id, v0, vtan):
    t = np.ascontiguousarray(tangent, dtype=np.float64)
    n, dim = t.shape
    F = np.zeros_like(t)
    lib().orc_self_propel(dim, n, np.ascontiguousarray(chain_id, dtype=np.int64), v0,
                          np.ascontiguousarray(vtan, dtype=np.float64), t, F)
    return F


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def bead_noise(dim, seed, orig, step):
    xi = np.zeros(2 * dim)
    lib().orc_bead_noise(dim, seed, orig, step, xi)
    return xi


def walls(pos, L, wca, mask, zone):
    """orc_walls: the wall forces on beads at pos."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, dim = pos.shape
    F = np.zeros_like(pos)
    lib().orc_walls(dim, n, pos, L, np.ascontiguousarray(wca, dtype=np.float64), int(mask),
                    float(zone), F)
    return F


def force_eval(sysd, pos, chain_id, rest_lengths, vtan, tangent):
    s = make_sys(sysd)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    T = np.array(tangent, dtype=np.float64)
    F = np.zeros_like(pos)
    lib().orc_force_eval(C.byref(s), pos, np.ascontiguousarray(chain_id, dtype=np.int64),
                         np.ascontiguousarray(rest_lengths, dtype=np.float64),
                         np.ascontiguousarray(vtan, dtype=np.float64), T, F, 0, None, None)
    return F, T


def cross_pair_forces(posA, posB, L, p):
    posA = np.ascontiguousarray(posA, dtype=np.float64)
    posB = np.ascontiguousarray(posB, dtype=np.float64)
    F = np.zeros_like(posA)
    lib().orc_cross_pair_forces(posA.shape[1], len(posA), posA.ctypes.data, len(posB),
                                posB.ctypes.data, float(L),
                                np.ascontiguousarray(p, dtype=np.float64).ctypes.data,
                                F.ctypes.data)
    return F


def multi_srk2_run(pops, cross, time_vals, step0=0, pair_mode=0, n_threads=1, grid_factor=2.0):
    """orc_multi_srk2_run. pops: list of dicts with sysd, pos, tangent,
    chain_id, rest_lengths, vtan, and optionally xi, seed, fixed, motor (a
    dict max_dist, imparts, k_on, k_off, reach, still; the per-motor bound
    bead is returned in d["motor_bead"]); cross: {(P, Q): wca 4-vector, or
    "motor" for the pushes of motor population Q on P}. pair_mode 1 scans
    each population's own pairs over Verlet lists and the other populations
    through cell grids, on n_threads threads (the CPU baseline). Returns a
    list of (pos, pos1, tangent, forces) per population and the status."""
    n = len(pops)
    assert 1 <= n <= 16
    arr = (OrcPop * n)()
    keep = []
    out = []
    for k, d in enumerate(pops):
        pos = np.array(d["pos"], dtype=np.float64)
        pos1 = pos.copy()
        T = np.array(d["tangent"], dtype=np.float64)
        F = np.zeros_like(pos)
        cid = np.ascontiguousarray(d["chain_id"], dtype=np.int64)
        rl = np.ascontiguousarray(d["rest_lengths"], dtype=np.float64)
        vt = np.ascontiguousarray(d["vtan"], dtype=np.float64)
        xi = d.get("xi")
        xi = None if xi is None else np.ascontiguousarray(xi, dtype=np.float64)
        keep += [cid, rl, vt, xi]
        arr[k].sys = make_sys(d["sysd"])
        arr[k].fixed = int(bool(d.get("fixed", False)))
        arr[k].pos, arr[k].pos1 = pos.ctypes.data, pos1.ctypes.data
        arr[k].tangent, arr[k].forces = T.ctypes.data, F.ctypes.data
        arr[k].chain_id, arr[k].rest_lengths = cid.ctypes.data, rl.ctypes.data
        arr[k].vtan = vt.ctypes.data
        arr[k].xi = xi.ctypes.data if xi is not None else None
        arr[k].seed = int(d.get("seed", 0)) & ((1 << 64) - 1)
        mo = d.get("motor")
        if mo is not None:
            mb = d.setdefault("motor_bead", np.full(len(pos), -1, dtype=np.int32))
            keep.append(mb)
            arr[k].motor = 1
            arr[k].motor_still = int(bool(mo.get("still", True)))
            arr[k].motor_max_dist_sq = float(mo["max_dist"]) ** 2
            arr[k].motor_imparts = float(mo["imparts"])
            arr[k].k_on, arr[k].k_off = float(mo["k_on"]), float(mo["k_off"])
            arr[k].motor_reach = float(mo.get("reach", mo["max_dist"]))
            arr[k].motor_bead = mb.ctypes.data
        out.append((pos, pos1, T, F))
    cr = np.zeros((n, n, 5))
    for (a, b), p in cross.items():
        if isinstance(p, str) and p == "motor":
            cr[a, b, 0] = 2.0
            continue
        cr[a, b, 0] = 1.0
        cr[a, b, 1:] = p
    st = np.zeros(4)
    tv = np.ascontiguousarray(time_vals, dtype=np.float64)
    lib().orc_multi_srk2_run(n, arr, cr.ctypes.data, tv.ctypes.data, st.ctypes.data, int(step0),
                             int(pair_mode), int(n_threads), float(grid_factor))
    return out, st


def rod_project(prev, F, xi, chain_id, rest_lengths, factor, L, dt, D, radius):
    """orc_rod_project: positions after one projected update."""
    prev = np.ascontiguousarray(prev, dtype=np.float64)
    n, dim = prev.shape
    rtn = np.zeros_like(prev)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    f = lib().orc_rod_project
    f.restype = None
    f.argtypes = [C.c_int, C.c_int, np.ctypeslib.ndpointer(np.float64),
                  np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.float64),
                  np.ctypeslib.ndpointer(np.float64), C.c_int, np.ctypeslib.ndpointer(np.int64),
                  np.ctypeslib.ndpointer(np.float64), C.c_double, C.c_double, C.c_double,
                  C.c_double, C.c_double]
    f(dim, n, rtn, prev, np.ascontiguousarray(F, dtype=np.float64), xi, xi.shape[1],
      np.ascontiguousarray(chain_id, dtype=np.int64),
      np.ascontiguousarray(rest_lengths, dtype=np.float64), factor, L, dt, D, radius)
    return rtn


def srk2_run(sysd, pos, tangent, chain_id, rest_lengths, vtan, time_vals, xi=None, seed=0,
             step0=0, pair_mode=0, n_threads=1, grid_factor=2.0):
    """Returns (pos, pos1, tangent, forces, status[dt, elapsed, steps, m])."""
    s = make_sys(sysd)
    pos = np.array(pos, dtype=np.float64)
    pos1 = pos.copy()
    T = np.array(tangent, dtype=np.float64)
    F = np.zeros_like(pos)
    st = np.zeros(4)
    xi_arr = None
    if xi is not None:
        xi_arr = np.ascontiguousarray(xi, dtype=np.float64)
    lib().orc_srk2_run(C.byref(s), pos, pos1, T, F,
                       np.ascontiguousarray(chain_id, dtype=np.int64),
                       np.ascontiguousarray(rest_lengths, dtype=np.float64),
                       np.ascontiguousarray(vtan, dtype=np.float64),
                       np.ascontiguousarray(time_vals, dtype=np.float64), st,
                       xi_arr.ctypes.data if xi_arr is not None else None, seed, step0,
                       pair_mode, n_threads, grid_factor)
    return pos, pos1, T, F, st
