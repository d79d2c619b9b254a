"""The device cell list is an exact stable sort (updates.py:502-583).

Index work is checked bit for bit: after a re-sort, gcr_replica_cell_list's
first-slot table and bead slots equal what a stable argsort of the beads'
cell keys gives on the host, for 2D and 3D replicas, alone and in a pool of
several (slots counted from each replica's first slot), at the initial
positions and again after the replicas have moved."""
import numpy as np
import pytest

from gcr_amd import engine as E
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu


def expected_cell_list(pos, L, nc, dim):
    """(cstart, slot) of a stable sort of the beads by cell key
    (z nc + y) nc + x with x = int(pos * (nc / L)) clamped into [0, nc)."""
    inv_h = np.float64(nc) / np.float64(L)
    c = np.clip((pos * inv_h).astype(np.int64), 0, nc - 1)
    key = c[:, 0] + nc * c[:, 1]
    if dim == 3:
        key = key + nc * nc * c[:, 2]
    order = np.argsort(key, kind="stable")
    slot = np.empty(len(pos), dtype=np.int64)
    slot[order] = np.arange(len(pos))
    counts = np.bincount(key, minlength=nc ** dim)
    cstart = np.concatenate([[0], np.cumsum(counts)])
    return cstart, slot


def check(pool, rid, w):
    pos = pool.download(rid, pos1=False, tangent=False, forces=False)["pos"]
    nc, cstart, slot = pool.cell_list(rid)
    want_cstart, want_slot = expected_cell_list(pos, w.phys.L, nc, w.phys.dim)
    assert np.array_equal(cstart, want_cstart), w.name
    assert np.array_equal(slot, want_slot), w.name
    return nc


def workloads(dim):
    if dim == 2:
        return [W.chains("cells2d_a", 2, 300, 10, phi=0.35, cutoff_factor=W.LJ_25_SIGMA, kappa=5.0,
                         vtan=0.2, T=0.05, dtmax=5e-4, steps=40, seed=71),
                W.chains("cells2d_b", 2, 900, 1, phi=0.3, T=0.01, Dr=0.1, steps=40, seed=72),
                W.chains("cells2d_c", 2, 50, 7, L=20.0, kappa=0.0, T=0.01, steps=40, seed=73)]
    return [W.chains("cells3d_a", 3, 200, 8, L=16.0, kappa=4.0, T=0.01, steps=40, seed=74),
            W.chains("cells3d_b", 3, 800, 1, L=12.0, T=0.01, Dr=0.1, steps=40, seed=75)]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("together", [False, True])
def test_cell_list_is_a_stable_sort(dim, together):
    ws = workloads(dim)
    groups = [ws] if together else [[w] for w in ws]
    for grp in groups:
        nb = sum(w.n for w in grp)
        pool = E.Pool(dim, E.GCR_FP64, bead_capacity=nb + 64,
                      cell_capacity=sum(E.replica_cells(w.phys) for w in grp) + 16,
                      chain_capacity=sum(w.phys.n_chains for w in grp) + 8,
                      replica_capacity=len(grp) + 1)
        try:
            pool.set_schedule(16, 8)
            rids = [pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths,
                             w.v_tan_angle) for w in grp]
            for w, r in zip(grp, rids):  # the re-sort of gcr_replica_add
                nc = check(pool, r, w)
                print(f"[cells] {w.name} n={w.n} nc={nc} together={together}", flush=True)
            for w, r in zip(grp, rids):
                tv = w.time_vals.copy()
                tv[7] = 40
                pool.begin(r, tv)
            pool.run()
            for w, r in zip(grp, rids):  # a re-sort at the moved positions
                pool.eval_forces(r)
                check(pool, r, w)
        finally:
            pool.close()
