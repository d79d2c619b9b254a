"""C-ABI checks that need no GPU: libgcr.so loads, exports every function
include/gcr.h declares, rejects bad arguments with an error code, and the
host-side initial placement and workload constants are right."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from gcr_amd import _lib
from gcr_amd import workloads as W


def header_functions(repo_root):
    text = open(os.path.join(repo_root, "include", "gcr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gcr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header(repo_root):
    lib = _lib.load()
    names = header_functions(repo_root)
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gcr_\w+)", out))
    assert set(names) <= exported
    assert set(_lib.EXPORTS) <= exported


def test_bad_arguments_return_error_codes():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.gcr_pool_create(4, 0, 100, 100, 10, 1, C.byref(h)) == -1  # dim
    assert lib.gcr_pool_create(2, 7, 100, 100, 10, 1, C.byref(h)) == -1  # precision
    assert lib.gcr_pool_create(2, 0, -1, 100, 10, 1, C.byref(h)) == -1   # capacity
    assert b"dim" in lib.gcr_last_error() or lib.gcr_last_error()
    assert lib.gcr_pool_set_schedule(None, 8, 4) == -1
    assert lib.gcr_replica_remove(None, 0) == -1
    assert lib.gcr_pool_destroy(None) == 0
    d = W.chains("t", 2, 3, 5, L=15.0).phys.desc()
    assert lib.gcr_replica_cells(2, C.byref(d)) == int(15.0 // (1.3)) ** 2


def test_init_chains_non_overlapping():
    from gcr_amd.engine import init_chains
    L = 30.0
    lens = np.array([7, 7, 3, 1, 12] * 8, dtype=np.int32)
    pos, tang = init_chains(2, L, lens, 0.5, min_dist=0.9, seed=3)
    chain = np.repeat(np.arange(len(lens)), lens)
    assert np.all((pos >= 0) & (pos < L))
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    other = chain[:, None] != chain[None, :]
    assert r[other].min() >= 0.9
    # bonds of length 0.5; tangent points from bead j+1 towards bead j
    b0 = 0
    for ln in lens:
        for j in range(ln - 1):
            t = pos[b0 + j] - pos[b0 + j + 1]
            t -= L * np.round(t / L)
            assert np.linalg.norm(t) == pytest.approx(0.5, rel=1e-9)
            assert np.allclose(t / 0.5, tang[b0 + j], atol=1e-9)
        b0 += ln
    # same seed, same placement
    pos2, _ = init_chains(2, L, lens, 0.5, min_dist=0.9, seed=3)
    assert np.array_equal(pos, pos2)


def test_workload_constants():
    """Box edges from area fraction (inverse of config.py:80-84) and
    max_dist_per_step of config.py:182-189 for the benchmark systems."""
    assert W.box_from_area_fraction(1000, 20, 0.5, 1.0, 0.3) == pytest.approx(35000 ** 0.5)
    assert W.box_from_area_fraction(20000, 20, 0.5, 1.0, 0.4) == pytest.approx(525000 ** 0.5)
    mdps, n_upd = W.grid_constants(0.5, 7, 1.0)
    assert (mdps, n_upd) == (0.0125, 40)
    mdps, _ = W.grid_constants(0.5, 20, W.LJ_25_SIGMA)
    assert mdps == pytest.approx(0.5 * W.LJ_25_SIGMA / 40)
    p = W.wca_params(1.0, 1.0, 0.0, W.LJ_25_SIGMA)
    assert p[3] == pytest.approx(2.5 / 2 ** (1 / 6)) and p[0] == 12.0
    w = W.chains_4x7(seed=1)
    assert w.n == 28 and w.phys.D == 0.01 and w.phys.Dr == 0.0075
    assert w.time_vals[0] == pytest.approx(1e-7) and w.time_vals[1] == 1e-3
    assert w.phys.spring[4] == 1.0 and w.phys.spring[1] == 0.0


def test_cell_grid_sizes():
    """gcr_replica_cells: with a pair force the cell edge is at least the
    cutoff plus the skin; without one (a motor population, say) the grid holds
    about four beads per cell, and tiny populations keep a single cell."""
    from gcr_amd import engine as E
    lib = _lib.load()

    def cells(**kw):
        p = E.Physics(D=0.0, Dr=0.0, **kw)
        return int(lib.gcr_replica_cells(p.dim, C.byref(p.desc())))

    L = 100.0
    wca = np.array([12.0, 12.0, 0.0, 2.5])
    nc = int(np.floor(L / (2.5 + 0.3)))
    assert cells(dim=2, n_beads=1000, n_chains=1000, L=L, wca=wca, skin=0.3) == nc * nc
    assert cells(dim=3, n_beads=1000, n_chains=1000, L=L, wca=wca, skin=0.3) == nc ** 3
    free = dict(n_chains=1, wca=np.array([0.0, 0.0, 0.0, 1.0]), has_wca=False, has_springs=False,
                has_propel=False, skin=0.3)
    n = 40000
    side = int(np.floor(L / np.sqrt(4.0 * L * L / n)))
    assert cells(dim=2, n_beads=n, L=L, **free) == side * side
    assert 3.0 <= n / side ** 2 <= 5.0
    side3 = int(np.floor(L / (4.0 * L ** 3 / n) ** (1 / 3)))
    assert cells(dim=3, n_beads=n, L=L, **free) == side3 ** 3
    assert cells(dim=2, n_beads=8, L=L, **free) == 1
