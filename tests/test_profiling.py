"""CPU checks of the rocprofv3 summary helpers behind the roofline figures:
the window between the two modularity passes, per-dispatch counter means and
the copy calibration that scales FETCH_SIZE / WRITE_SIZE to bytes."""
import csv
import os

from hit_leiden_amd import profiling as P


def write_csv(path, rows, fields):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def counters(d, counter, samples):
    """samples: [(kernel, start, value)] -> a counter_collection.csv"""
    rows = [{"Kernel_Name": k, "Start_Timestamp": t, "End_Timestamp": t + 1, "Counter_Name": counter,
             "Counter_Value": v} for k, t, v in samples]
    write_csv(os.path.join(d, "run_counter_collection.csv"), rows,
              ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Counter_Name", "Counter_Value"])


def test_window_and_traffic(tmp_path):
    fetch, write = str(tmp_path / "fetch"), str(tmp_path / "write")
    # a load-time dispatch before the first k_intra is left out of the window
    seq = [("k_prop2_tab<hl::MovePolicyT, 32u, 64u>(args)", 1, 900.0), ("k_intra(args)", 2, 1.0),
           ("k_prop2_tab<hl::MovePolicyT, 32u, 64u>(args)", 3, 100.0),
           ("k_prop2_tab<hl::RefinePolicyT, 32u, 64u>(args)", 4, 300.0), ("k_dec_small(args)", 5, 7.0),
           ("k_intra(args)", 6, 1.0)]
    counters(fetch, "FETCH_SIZE", seq)
    counters(write, "WRITE_SIZE", [(k, t, v / 10) for k, t, v in seq])
    rx = "k_prop2_tab<hl::(Move|Refine)PolicyT, 32u?, 64u?>"
    m = P.window_counter_means(fetch, rx)
    assert m["FETCH_SIZE"] == (2, 200.0)
    e = P.traffic_entry(fetch, write, rx, 2.0, 1.0, "k_prop2_tab")
    assert e["dispatches"] == 2
    assert e["bytes_per_launch"] == int(round(200.0 * 1024 * 2.0 + 20.0 * 1024 * 1.0))


def test_copy_calibration(tmp_path):
    cf, cw = str(tmp_path / "cf"), str(tmp_path / "cw")
    counters(cf, "FETCH_SIZE", [("__amd_rocclr_copyBuffer", 1, 131072.0), ("__amd_rocclr_copyBuffer", 2, 131072.0)])
    counters(cw, "WRITE_SIZE", [("__amd_rocclr_copyBuffer", 1, 262144.0)])
    cal = P.calibration(cf, cw, 1 << 28)
    assert abs(cal["fetch_factor"] - 2.0) < 1e-9 and abs(cal["write_factor"] - 1.0) < 1e-9


def test_kernel_stats_and_summary(tmp_path):
    d = str(tmp_path / "trace")
    write_csv(os.path.join(d, "run_kernel_stats.csv"),
              [{"Name": "k_dec_small(args)", "Calls": 4, "TotalDurationNs": 80000, "AverageNs": 20000,
                "Percentage": 60.0},
               {"Name": "k_prop2_first<hl::MovePolicyT>(args)", "Calls": 2, "TotalDurationNs": 20000,
                "AverageNs": 10000, "Percentage": 40.0}],
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"])
    rows = [{"Kernel_Name": n, "Start_Timestamp": a, "End_Timestamp": b}
            for n, a, b in [("k_intra(x)", 0, 10), ("k_dec_small(x)", 20, 2020), ("k_dec_small(x)", 3000, 7000),
                            ("k_intra(x)", 8000, 8010)]]
    write_csv(os.path.join(d, "run_kernel_trace.csv"), rows, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
    assert P.kernel_avg_us(d, "k_dec_small") == (4, 20.0)
    win, span = P.window_kernel_stats(d)
    assert [(w[0], w[1]) for w in win] == [("k_dec_small(x)", 2)]
    assert abs(span - 6980 / 1e6) < 1e-12  # ms, first start to last end inside the window
    text = P.summary_text(d, "cmd", pattern="k_dec_small")
    assert "kernels matching /k_dec_small/: 4 calls, call-weighted average 20.00 us" in text
    assert "in this window: 2 calls, call-weighted average 3.00 us" in text
