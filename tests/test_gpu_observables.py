"""The observables series kept on the device (gcr_replica_sample /
gcr_replica_series) against the host observables on downloaded arrays
(gcr_amd.observables, pinned by closed forms in tests/test_observables.py):
polar and nematic order and the director of the chains' mean directions
(analysis.py:294-328, output_methods.py:492-510), the mean square
displacement of the chain centres of mass (host: every bead unwrapped sample
to sample, then the chain means), the simulated time and the step count."""
import numpy as np
import pytest

from gcr_amd import engine as E
from gcr_amd import observables as OB
from gcr_amd import workloads as W

pytestmark = pytest.mark.gpu


def workload(kind):
    if kind == "chains2d":
        return W.chains("s2", 2, 300, 10, phi=0.3, kappa=10.0, vtan=0.3, T=0.01, steps=50, seed=41)
    if kind == "abp2d":
        return W.chains("sa", 2, 2000, 1, phi=0.3, T=0.05, Dr=0.1, steps=50, seed=42)
    return W.chains("s3", 3, 100, 6, L=14.0, kappa=4.0, T=0.01, steps=50, seed=43)


@pytest.mark.parametrize("precision", [E.GCR_FP64, E.GCR_MIXED])
@pytest.mark.parametrize("kind", ["chains2d", "abp2d", "chains3d"])
def test_device_series_matches_host(kind, precision):
    w = workload(kind)
    pool = E.Pool(w.phys.dim, precision, bead_capacity=w.n,
                  cell_capacity=E.replica_cells(w.phys) + 1, chain_capacity=w.phys.n_chains,
                  replica_capacity=1)
    try:
        pool.set_schedule(16, 8)
        rid = pool.add(w.phys, w.pos, w.tangent, w.chain_id, w.rest_lengths, w.v_tan_angle)
        tv = w.time_vals.copy()
        tv[7] = 40
        pool.begin(rid, tv)
        pool.run()  # some motion before the origin
        t_host = pool.status(rid)["elapsed"]
        pool.sample(rid, reset=True)
        unw = OB.Unwrapper(pool.download(rid, pos1=False, forces=False)["pos"], w.phys.L)
        rows = []
        for k in range(12):
            tv[0] = pool.status(rid)["dt"]
            tv[7] = 25
            pool.begin(rid, tv)
            pool.run()
            st = pool.status(rid)
            t_host += st["elapsed"]
            pool.sample(rid)
            d = pool.download(rid, pos1=False, forces=False)
            unw.add(d["pos"])
            phi = OB.chain_mean_angle(d["tangent"], w.chain_id)
            S, director = OB.nematic_order(phi)
            rows.append((t_host, OB.polar_order(phi), S, director,
                         unw.chain_com_msd(w.chain_id), st["step_global"]))
        ser = pool.series(rid)
    finally:
        pool.close()
    host = np.array(rows)
    print(f"[series] {kind} precision={precision} msd {ser['msd'][-1]:.4e} "
          f"polar {ser['polar'][-1]:.4f}", flush=True)
    assert len(ser["t"]) == 13 and ser["msd"][0] == 0.0
    dev = np.stack([ser["t"], ser["polar"], ser["nematic"], ser["director"], ser["msd"],
                    ser["step"]], axis=1)[1:]
    assert np.allclose(dev[:, 0], host[:, 0], rtol=1e-12, atol=0)
    assert np.max(np.abs(dev[:, 1] - host[:, 1])) < 1e-12
    assert np.max(np.abs(dev[:, 2] - host[:, 2])) < 1e-12
    assert np.max(np.abs(dev[:, 3] - host[:, 3])) < 1e-9
    assert np.max(np.abs(dev[:, 4] - host[:, 4]) / (host[:, 4] + 1e-12)) < 1e-9
    assert np.array_equal(dev[:, 5], host[:, 5])
    assert dev[-1, 4] > 0.0
