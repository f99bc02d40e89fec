"""Writes hete_media.tab, the DIST_HETE1 grid of this case (seeded; run once)."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "..", ".."))
from sem2dpack_amd import workloads  # noqa: E402

g = workloads.layered_random_grid((0.0, 8e3), (0.0, 4e3), dx=100.0, corr_len=600.0, seed=7)
workloads.write_hete1(os.path.join(os.path.dirname(os.path.abspath(__file__)), "hete_media.tab"), *g)
