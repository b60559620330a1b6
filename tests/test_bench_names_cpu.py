"""bench.py names its timed kernel groups' kernels to find their HBM
traffic in the committed FETCH_SIZE/WRITE_SIZE summary; every group it
reports must resolve there to exactly one kernel of each precision (a
template argument added to a kernel would otherwise leave the bench line's
roofline `traffic` empty)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.mark.parametrize("vtype", ["double", "float"])
def test_traffic_names_resolve(vtype):
    import bench
    path = os.path.join(ROOT, "profiles", "lagoon50k_pmc_traffic.md")
    for group in bench.KERNEL_NAMES:
        t = bench._measured_traffic(path, group, vtype)
        assert t is not None and t > 0, (group, vtype)
