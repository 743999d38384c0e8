"""bench.py's reading of the committed profile: the RK stage's HBM traffic
per launch comes from the FETCH_SIZE/WRITE_SIZE table only, averaged over
the dispatches of the variant's instances (other tables in the same file
have rows for the same kernels)."""
import os

import bench


def test_rk_traffic_reads_the_traffic_table_only():
    path = os.path.join(os.path.dirname(bench.__file__), "profiles",
                        "mhd_step_turbulence_512cubed_f64.md")
    rows = []
    in_traffic = False
    for line in open(path):
        if line.startswith("| kernel |"):
            in_traffic = "traffic MB" in line
            continue
        if in_traffic and line.startswith("| `rk_stage_kernel_march<double, 8"):
            cols = [c.strip() for c in line.split("|")]
            rows.append((int(cols[2]), float(cols[7])))
    assert rows, "the profile has no traffic rows for the z-march RK stage"
    expect = sum(n * mb for n, mb in rows) / sum(n for n, _ in rows) * 1e6
    got = bench.rk_traffic(512, "f64", 5)
    assert abs(got - expect) <= 1e-9 * expect
    # the traffic of an RK stage is tens of GB at 512^3, not a clock or a count
    assert 2e10 < got < 6e10
    assert bench.rk_traffic(96, "f64", 5) is None
