"""State snapshots in the reference's `.dat` format.

save_state writes what output_methods.save_state (output_methods.py:420-485)
writes for a chain population: a two-line `# ` header
(`filenumber N stepnum S t T class populations.<Class>` and the column
names) and one row per bead, `chain_id inherited_label x y [z] phi [theta]
where_active` with the formats `%d %d %.8e ... %d` of np.savetxt, under the
file name `<prefix>_<pop name>[_<zero-padded filenumber>].dat`. The
orientations are those of populations.tangents_to_orientations
(populations.py:559-562): phi = atan2(t_y, t_x), theta = arccos(t_z) in 3D.
read_state parses such a file the way analysis.recall (analysis.py:37-104)
does and returns the arrays, with tangents rebuilt from the orientations.

The arrays come from Pool.download / System.download (positions and
tangent_list after a call, in the reference's bead order). The table is
formatted by libgcr's gcr_write_table on several host threads, byte for byte
what np.savetxt writes (np.savetxt alone takes about half a second for 400k
beads, several times the GPU time between two snapshots).
"""
import os

import numpy as np

from . import _lib

CHAIN_CLASS = "populations.Equal_Size_Chain_Collection"


def zero_padded(number, n_digits):
    """output_methods.zero_padded_string: zeros on the left, n_digits wide."""
    return ("0" * n_digits)[:-len(str(number))] + str(number)


def state_filename(prefix, pop_name, filenumber, n_digits=8, most_recent_only=False):
    name = prefix + "_" + pop_name
    if not most_recent_only:
        name += "_" + zero_padded(filenumber, n_digits)
    return name + ".dat"


def orientations(tangent):
    tangent = np.asarray(tangent, dtype=np.float64)
    phi = np.arctan2(tangent[:, 1], tangent[:, 0])
    if tangent.shape[1] == 3:
        return np.column_stack((phi, np.arccos(tangent[:, 2])))
    return phi[:, None]


def save_state(prefix, pop_name, filenumber, stepnum, t, pos, tangent, chain_id,
               inherited_label=None, where_active=None, class_name=CHAIN_CLASS, n_digits=8,
               most_recent_only=False):
    """Write one population's snapshot; returns the file name."""
    pos = np.asarray(pos, dtype=np.float64)
    n, dim = pos.shape
    chain_id = np.asarray(chain_id)
    inherited_label = chain_id.copy() if inherited_label is None else np.asarray(inherited_label)
    where_active = -np.ones(n, dtype=np.int64) if where_active is None else np.asarray(where_active)
    ori = orientations(tangent)
    fname = state_filename(prefix, pop_name, filenumber, n_digits, most_recent_only)
    head = f"filenumber {filenumber} stepnum {stepnum} t {t:.5f} class {class_name}\n"
    if dim == 2:
        cols = (chain_id, inherited_label, pos[:, 0], pos[:, 1], ori[:, 0], where_active)
        fmt = ["%d", "%d", "%.8e", "%.8e", "%.8e", "%d"]
        head += "chain_id inherited_label x y phi where_active inherited_label"
    else:
        cols = (chain_id, inherited_label, pos[:, 0], pos[:, 1], pos[:, 2], ori[:, 0], ori[:, 1],
                where_active)
        fmt = ["%d", "%d", "%.8e", "%.8e", "%.8e", "%.8e", "%.8e", "%d"]
        head += "chain_id inherited_label x y z phi theta where_active inherited_label"
    table = np.ascontiguousarray(np.column_stack(cols).T)  # [column][row], float64 as savetxt's
    is_int = np.array([f == "%d" for f in fmt], dtype=np.int32)
    threads = min(16, os.cpu_count() or 1)
    _lib.check(_lib.load().gcr_write_table(fname.encode(), head.encode(), n, len(fmt),
                                           table.ctypes.data, is_int.ctypes.data, threads))
    return fname


def read_state(fname):
    """The arrays analysis.recall rebuilds from a `.dat` file."""
    with open(fname) as f:
        lines = f.readlines()
    headers = lines[1].split("#")[1].split()
    data = lines[2:]
    first = lines[0].split()[1:]
    out = dict(class_name=first[first.index("class") + 1],
               stepnum=int(first[first.index("stepnum") + 1]),
               filenumber=int(first[first.index("filenumber") + 1]),
               t=float(first[first.index("t") + 1]))

    def column(name, conv):
        k = headers.index(name)
        return np.array([conv(line.split(" ")[k]) for line in data])

    out["chain_id"] = column("chain_id", lambda v: int(float(v)))
    out["inherited_label"] = column("inherited_label", lambda v: int(float(v)))
    # by header name: analysis.recall reads column 5, which is where_active only in 2D
    out["where_active"] = column("where_active", lambda v: int(float(v)))
    out["pos"] = np.column_stack([column(a, float) for a in ("x", "y", "z") if a in headers])
    ori = np.column_stack([column(a, float) for a in ("phi", "theta") if a in headers])
    if ori.shape[1] == 1:
        out["tangent"] = np.column_stack((np.cos(ori[:, 0]), np.sin(ori[:, 0])))
    else:
        s = np.sin(ori[:, 1])
        out["tangent"] = np.column_stack((s * np.cos(ori[:, 0]), s * np.sin(ori[:, 0]),
                                          np.cos(ori[:, 1])))
    out["orientations"] = ori
    return out
