"""gcr_eval_forces_noalloc (the force evaluation that allocates no device
memory) against gcr_eval_forces on the same systems, through the C-ABI: two
calls in a row, the second on a smaller system, so the second call runs on
the array the first one left behind. The bar is that of the oracle parity
(forces within 1e-11 of the largest force, tangents 1e-13); both entry points
run the same kernels, so the results agree to the rounding of the chiral
angle's cosine and sine (taken on the device here, on the host there).

Device buffers come from hipMalloc of the HIP runtime libgcr.so is linked
against, called through ctypes."""
import ctypes as C

import numpy as np
import pytest

from gcr_amd import _lib
from gcr_amd import engine as E
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    _lib.load()  # libamdhip64.so.7 is then the runtime libgcr.so uses
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


class DevArrays:
    """Device copies of host arrays (freed by close)."""

    def __init__(self, h):
        self.h = h
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def noalloc_forces(w, precision):
    h = hip()
    L = _lib.load()
    dev = DevArrays(h)
    try:
        n, dim = w.n, w.phys.dim
        pos = dev.put(np.asarray(w.pos, dtype=np.float64))
        tang = dev.put(np.asarray(w.tangent, dtype=np.float64))
        cid = dev.put(np.asarray(w.chain_id, dtype=np.int64))
        rest = dev.put(np.asarray(w.rest_lengths, dtype=np.float64))
        vt = dev.put(np.asarray(w.v_tan_angle, dtype=np.float64))
        F = dev.empty(n * dim * 8)
        T = dev.empty(n * dim * 8)
        d = w.phys.desc()
        _lib.check(L.gcr_eval_forces_noalloc(C.byref(d), dim, precision, pos, tang, cid, rest, vt,
                                             F, T, None))
        assert h.hipDeviceSynchronize() == 0
        return dev.get(F, (n, dim)), dev.get(T, (n, dim))
    finally:
        dev.close()


def pool_forces(w, precision):
    s = E.System(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle,
                 precision=precision)
    try:
        return s.eval_forces()
    finally:
        s.close()


def pairs():
    rng = np.random.default_rng(3)
    big2 = W.chains("na_big2d", 2, 400, 20, phi=0.4, cutoff_factor=W.LJ_25_SIGMA, kappa=5.0,
                    vtan=rng.uniform(-0.5, 0.5, 400), T=0.05, dtmax=5e-4, steps=10, seed=61)
    small2 = W.chains("na_small2d", 2, 40, 10, L=22.0, kappa=3.0, shift=0.1, n_nearby=-1,
                      vtan=0.2, T=0.01, steps=10, seed=62)
    big3 = W.chains("na_big3d", 3, 300, 8, L=16.0, kappa=4.0, T=0.01, steps=10, seed=63)
    small3 = W.chains("na_small3d", 3, 300, 1, L=10.0, T=0.01, Dr=0.1, steps=10, seed=64)
    return {"2d": (big2, small2), "3d": (big3, small3)}


PAIRS = pairs()


@pytest.mark.parametrize("precision", [E.GCR_FP64, E.GCR_MIXED])
@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_noalloc_matches_eval_forces_twice(kind, precision):
    for w in PAIRS[kind]:  # the second call on a smaller system
        print(f"[noalloc] {w.name} n={w.n} dim={w.phys.dim} precision={precision}", flush=True)
        F_n, T_n = noalloc_forces(w, precision)
        F_p, T_p = pool_forces(w, precision)
        scale = max(1.0, float(np.max(np.abs(F_p))))
        assert np.all(np.isfinite(F_n))
        assert np.max(np.abs(F_n - F_p)) <= 1e-11 * scale, w.name
        assert np.max(np.abs(T_n - T_p)) <= 1e-13, w.name


def test_noalloc_rejects_sizes_beyond_its_array():
    """More beads than GCR_NOALLOC_MAX_BEADS: GCR_ERR_CAPACITY before any
    device work (null device pointers are never touched)."""
    L = _lib.load()
    w = W.chains("na_too_big", 2, 1700, 20, phi=0.3, T=0.01, steps=1, seed=65)
    assert w.n > 32768
    d = w.phys.desc()
    one = C.c_void_p(8)  # never dereferenced: the size check comes first
    rc = L.gcr_eval_forces_noalloc(C.byref(d), 2, E.GCR_FP64, one, one, one, one, one, one,
                                   one, None)
    assert rc == -3
    assert b"GCR_NOALLOC_MAX_BEADS" in L.gcr_last_error()


def test_noalloc_flags_chain_ids_it_cannot_use():
    """Chain ids that are not contiguous runs of 0..n_chains-1 (here bead 5
    of chain 0 labelled 1, which splits chains 0 and 1 into two runs each)
    give NaN forces instead of reading out of bounds."""
    import copy
    w = PAIRS["2d"][1]
    bad = w.chain_id.copy()
    bad[5] = 1
    w2 = copy.copy(w)
    w2.chain_id = bad
    F, T = noalloc_forces(w2, E.GCR_FP64)
    assert np.all(np.isnan(F)) and np.all(np.isnan(T))
