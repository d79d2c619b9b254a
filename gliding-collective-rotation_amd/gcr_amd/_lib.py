"""ctypes binding of libgcr.so (include/gcr.h).

There is no CPU fallback: if the library is missing or cannot find a HIP
device the calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcr.so")

GCR_FP64 = 0
GCR_MIXED = 1


class GcrError(RuntimeError):
    """An error code of libgcr (rc: GCR_ERR_ARG -1, GCR_ERR_HIP -2,
    GCR_ERR_CAPACITY -3, GCR_ERR_STATE -4) with its message."""

    def __init__(self, msg, rc=None):
        super().__init__(msg)
        self.rc = rc


GCR_ERR_CAPACITY = -3


class ReplicaDesc(C.Structure):
    """gcr_replica_desc of include/gcr.h."""

    _fields_ = [
        ("n_beads", C.c_int32),
        ("n_chains", C.c_int32),
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
        ("skin", C.c_double),
        ("seed", C.c_uint64),
        ("rod", C.c_double),
        ("srk_order", C.c_int32),
        ("walls", C.c_int32),
        ("wall_zone", C.c_double),
        ("wall_cutoff", C.c_double),
        ("sliding_maxdist", C.c_double),
        ("sliding_v0", C.c_double),
        ("motor", C.c_int32),
        ("motor_still", C.c_int32),
        ("motor_max_dist", C.c_double),
        ("motor_imparts", C.c_double),
        ("k_on", C.c_double),
        ("k_off", C.c_double),
        ("motor_reach", C.c_double),
    ]


EXPORTS = [
    "gcr_last_error", "gcr_device_info", "gcr_grid_barrier_us", "gcr_create", "gcr_upload", "gcr_run", "gcr_download",
    "gcr_eval_forces", "gcr_destroy", "gcr_pool_create", "gcr_pool_set_schedule",
    "gcr_replica_bytes", "gcr_replica_cells", "gcr_replica_add", "gcr_replica_remove",
    "gcr_replica_upload", "gcr_replica_begin", "gcr_pool_run", "gcr_replica_status",
    "gcr_replica_download", "gcr_replica_eval_forces", "gcr_pool_time_kernels",
    "gcr_pool_destroy", "gcr_pool_set_tile", "gcr_pool_set_tile_shape", "gcr_pool_set_tile_threads",
    "gcr_init_chains",
    "gcr_replica_order", "gcr_replica_set_physics", "gcr_pool_group", "gcr_replica_set_chunks",
    "gcr_replica_where_active", "gcr_write_table", "gcr_eval_forces_noalloc",
    "gcr_replica_cell_list", "gcr_pool_graph_stats", "gcr_replica_sample", "gcr_replica_series",
]

_lib = None


def load():
    """Load libgcr.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GcrError(
            f"{LIB_PATH} is missing: build it with `python -m gcr_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    i32 = C.c_int32
    i64 = C.c_int64
    L.gcr_last_error.restype = C.c_char_p
    L.gcr_last_error.argtypes = []
    L.gcr_device_info.argtypes = [C.c_char_p, i32, C.POINTER(i32), C.POINTER(i64)]
    L.gcr_grid_barrier_us.argtypes = [i32, i32, C.POINTER(C.c_double)]
    L.gcr_pool_create.argtypes = [i32, i32, i64, i64, i64, i32, C.POINTER(vp)]
    L.gcr_pool_set_schedule.argtypes = [vp, i32, i32]
    L.gcr_replica_bytes.restype = i64
    L.gcr_replica_bytes.argtypes = [i32, i64, i64, i64]
    L.gcr_replica_cells.restype = i64
    L.gcr_replica_cells.argtypes = [i32, C.POINTER(ReplicaDesc)]
    L.gcr_replica_add.argtypes = [vp, C.POINTER(ReplicaDesc), vp, vp, vp, vp, vp, vp,
                                  C.POINTER(i32)]
    L.gcr_replica_remove.argtypes = [vp, i32]
    L.gcr_replica_upload.argtypes = [vp, i32, vp, vp]
    L.gcr_replica_begin.argtypes = [vp, i32, vp, vp, i64]
    L.gcr_pool_run.argtypes = [vp, i32, C.POINTER(i32)]
    L.gcr_replica_status.argtypes = [vp, i32, dp]
    L.gcr_replica_download.argtypes = [vp, i32, vp, vp, vp, vp]
    L.gcr_replica_eval_forces.argtypes = [vp, i32, vp, vp]
    L.gcr_pool_time_kernels.argtypes = [vp, i32, dp]
    L.gcr_pool_destroy.argtypes = [vp]
    L.gcr_pool_set_tile.argtypes = [vp, i32]
    L.gcr_pool_set_tile_shape.argtypes = [vp, i32, i32]
    L.gcr_pool_set_tile_threads.argtypes = [vp, i32]
    L.gcr_replica_order.argtypes = [vp, i32, dp]
    L.gcr_replica_set_physics.argtypes = [vp, i32, C.POINTER(ReplicaDesc), vp, vp]
    L.gcr_pool_group.argtypes = [vp, vp, i32, vp, vp]
    L.gcr_replica_set_chunks.argtypes = [vp, i32, vp]
    L.gcr_replica_where_active.argtypes = [vp, i32, vp]
    L.gcr_pool_graph_stats.argtypes = [vp, dp]
    L.gcr_replica_sample.argtypes = [vp, i32, i32]
    L.gcr_replica_series.argtypes = [vp, i32, vp, i32, C.POINTER(i32)]
    L.gcr_replica_cell_list.argtypes = [vp, i32, C.POINTER(i32), vp, vp]
    L.gcr_write_table.argtypes = [C.c_char_p, C.c_char_p, i64, i32, vp, vp, i32]
    L.gcr_create.argtypes = [C.POINTER(ReplicaDesc), i32, i32, vp, vp, vp, vp, vp, C.POINTER(vp)]
    L.gcr_upload.argtypes = [vp, vp, vp]
    L.gcr_run.argtypes = [vp, vp, vp, vp, i64]
    L.gcr_download.argtypes = [vp, vp, vp, vp, vp]
    L.gcr_eval_forces.argtypes = [vp, vp, vp]
    L.gcr_destroy.argtypes = [vp]
    L.gcr_eval_forces_noalloc.argtypes = [C.POINTER(ReplicaDesc), i32, i32, vp, vp, vp, vp, vp,
                                          vp, vp, vp]
    for name in EXPORTS:
        if name not in ("gcr_last_error", "gcr_replica_bytes", "gcr_replica_cells"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load().gcr_last_error().decode(errors="replace")
        raise GcrError(f"libgcr error {rc}: {msg}", rc)


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def device_info():
    L = load()
    name = C.create_string_buffer(256)
    ncu = C.c_int32()
    mem = C.c_int64()
    check(L.gcr_device_info(name, 256, C.byref(ncu), C.byref(mem)))
    return name.value.decode(), ncu.value, mem.value


def grid_barrier_us(blocks_per_cu=4, rounds=200):
    """Microseconds per grid-wide barrier of a cooperative launch with
    blocks_per_cu workgroups of 256 threads per CU (gcr_grid_barrier_us), and
    the workgroups it used."""
    L = load()
    out = (C.c_double * 2)()
    check(L.gcr_grid_barrier_us(int(blocks_per_cu), int(rounds), out))
    return float(out[0]), int(out[1])
