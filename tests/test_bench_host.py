"""bench.py's host-side pieces that need no GPU."""

import bench


def test_traffic_comes_from_the_committed_profile():
    traffic, kernel, src = bench.measured_traffic()
    assert src.startswith("profiles/") and traffic is not None and kernel.startswith("lbm::k_step<")
    # within 10% of the algorithmic 152 B per cell update of the workload
    assert abs(traffic / (152.0 * bench.NX * bench.NY * bench.NZ) - 1.0) < 0.1
