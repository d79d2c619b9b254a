"""MI355X Brownian-dynamics engine for active bead-spring chains (gcr_amd).

The hot path of the reference (WCA/LJ pair forces, Hookean bond + bending
forces, chiral self-propulsion, cell lists, SRK2 step) runs as hand-written
HIP kernels for gfx950 in libgcr.so, reached through the C-ABI declared in
include/gcr.h.
"""
from ._lib import GCR_FP64, GCR_MIXED, GcrError  # noqa: F401
