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


def test_trace_window_cuts_the_trace_at_the_markers(tmp_path, capsys):
    """--trace-window: kernels strictly between marker FIRST and FIRST+1,
    per step; the wall time runs from the first marker's end to the
    second's start, whatever order the CSV lists the dispatches in."""
    import csv
    rows = [("void misohip::before(int)", 0, 50),
            ("void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<short>, "
             "std::array<char*, 1ul> >(int)", 100, 110),
            ("void misohip::rk(int)", 1_000_110, 3_000_110),
            ("void misohip::walk(int)", 3_000_110, 4_000_110),
            ("void misohip::rk(int)", 5_000_110, 7_000_110),
            ("void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<short>, "
             "std::array<char*, 1ul> >(int)", 8_000_110, 8_000_120),
            ("void misohip::after(int)", 9_000_000, 9_500_000)]
    path = tmp_path / "trace.csv"
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for r in reversed(rows):
            w.writerow(r)
    bench.trace_window(str(path), 1, 2)
    out = capsys.readouterr().out
    lines = {ln.split("|")[1].strip(): ln for ln in out.splitlines() if ln.startswith("| `")}
    assert set(lines) == {"`rk`", "`walk`"}
    rk = [c.strip() for c in lines["`rk`"].split("|")]
    assert (rk[2], rk[3]) == ("1.0", "2.000")        # 2 launches / 2 steps, 4 ms / 2 steps
    walk = [c.strip() for c in lines["`walk`"].split("|")]
    assert (walk[2], walk[3]) == ("0.5", "0.500")
    assert "window: 4.000 ms per step wall, 2.500 ms per step of kernel time" in out
