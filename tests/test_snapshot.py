"""The `.dat` snapshot writer against the format of output_methods.save_state
(output_methods.py:420-485) and the parsing of analysis.recall
(analysis.py:37-104). The expected lines below are written out by hand from
that format: header `# filenumber N stepnum S t T(%.5f) class <module.Class>`,
column line, rows `%d %d %.8e %.8e %.8e %d` (2D)."""
import numpy as np

from gcr_amd import snapshot as SN


def test_file_name_zero_padding():
    assert SN.state_filename("Results/run", "beads", 7, n_digits=8) == "Results/run_beads_00000007.dat"
    assert SN.state_filename("r", "beads", 123456789, n_digits=8) == "r_beads_123456789.dat"
    assert SN.state_filename("r", "beads", 5, most_recent_only=True) == "r_beads.dat"


def test_2d_lines_exact(tmp_path):
    pos = np.array([[1.0, 2.0], [3.5, 0.25]])
    tangent = np.array([[np.sqrt(0.5), np.sqrt(0.5)], [-1.0, 0.0]])
    fname = SN.save_state(str(tmp_path / "run"), "beads", 3, 1000, 0.5, pos, tangent,
                          chain_id=np.array([0, 0]))
    lines = open(fname).read().splitlines()
    assert lines[0] == "# filenumber 3 stepnum 1000 t 0.50000 class populations.Equal_Size_Chain_Collection"
    assert lines[1] == "# chain_id inherited_label x y phi where_active inherited_label"
    assert lines[2] == "0 0 1.00000000e+00 2.00000000e+00 7.85398163e-01 -1"
    assert lines[3] == "0 0 3.50000000e+00 2.50000000e-01 3.14159265e+00 -1"


def test_round_trip_2d_and_3d(tmp_path):
    rng = np.random.default_rng(0)
    for dim in (2, 3):
        n = 40
        pos = rng.uniform(0, 20, (n, dim))
        t = rng.normal(size=(n, dim))
        t /= np.linalg.norm(t, axis=1)[:, None]
        cid = np.repeat(np.arange(8), 5)
        fname = SN.save_state(str(tmp_path / f"r{dim}"), "beads", 12, 34, 1.25, pos, t, cid)
        back = SN.read_state(fname)
        assert back["stepnum"] == 34 and back["filenumber"] == 12 and back["t"] == 1.25
        assert back["class_name"] == SN.CHAIN_CLASS
        assert np.array_equal(back["chain_id"], cid) and np.array_equal(back["inherited_label"], cid)
        assert np.all(back["where_active"] == -1)
        assert np.allclose(back["pos"], pos, rtol=1e-8, atol=0)
        assert np.allclose(back["tangent"], t, atol=1e-7)


def test_native_table_equals_numpy_savetxt(tmp_path):
    """gcr_write_table (the writer's threaded formatter) writes byte for byte
    what np.savetxt writes for the snapshot's columns and formats: 2D and 3D,
    50k rows over several threads, values from 1e-300 to 1e300 of either
    sign, signed zeros, integral columns with negatives."""
    rng = np.random.default_rng(1)
    for dim in (2, 3):
        n = 50000
        mag = 10.0 ** rng.uniform(-300, 300, (n, dim + dim - 1))
        vals = np.where(rng.random(mag.shape) < 0.5, -mag, mag)
        vals[:50] = 0.0
        vals[50:100] = -0.0
        vals[100:200] = rng.uniform(0, 20, (100, vals.shape[1]))
        pos, ori = vals[:, :dim], vals[:, dim:]
        cid = rng.integers(-5, 10 ** 9, n)
        wa = rng.integers(-3, 4, n)
        cols = [cid, cid[::-1].copy(), *pos.T, *ori.T, wa]
        fmt = ["%d", "%d"] + ["%.8e"] * (len(cols) - 3) + ["%d"]
        head = "filenumber 1 stepnum 2 t 3.00000 class X\nchain_id inherited_label cols"
        ref = tmp_path / f"ref{dim}.dat"
        np.savetxt(ref, np.column_stack(cols), fmt=fmt, header=head)
        out = tmp_path / f"out{dim}.dat"
        table = np.ascontiguousarray(np.column_stack(cols).T)
        is_int = np.array([f == "%d" for f in fmt], dtype=np.int32)
        SN._lib.check(SN._lib.load().gcr_write_table(str(out).encode(), head.encode(), n,
                                                     len(fmt), table.ctypes.data,
                                                     is_int.ctypes.data, 7))
        assert out.read_bytes() == ref.read_bytes()


def test_write_error_names_the_file(tmp_path):
    import pytest
    from gcr_amd import _lib
    with pytest.raises(_lib.GcrError, match="cannot open .*missing"):
        SN.save_state(str(tmp_path / "missing" / "run"), "beads", 1, 1, 0.1, np.zeros((3, 2)),
                      np.ones((3, 2)), np.zeros(3, dtype=np.int64))
