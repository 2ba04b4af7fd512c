"""The profile-window summary of bench.py: a rocprofv3 kernel trace or counter
collection is cut to the launches between the empty kernels that switching
profiling on (a burst) and off (one launch) issues."""
import csv
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _names():
    # before the window, the calibration burst, the window, the end marker, after
    return (["k_sweep<32, float>"] * 2 + ["k_empty"] * 3
            + ["k_sweep<32, float>", "k_pairs", "k_sweep<32, float>", "k_sweep<32, float>"]
            + ["k_empty"] + ["k_sweep<32, float>"] * 2)


def test_trace_window(tmp_path):
    p = tmp_path / "kernel_trace.csv"
    durs = [100, 100, 1, 1, 1, 10_000, 50_000, 2_000, 14_000, 1, 100, 100]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        t = 0
        for name, d in zip(_names(), durs):
            w.writerow({"Kernel_Name": "void (anonymous namespace)::" + name + "(SweepArgs<float>)",
                        "Start_Timestamp": t, "End_Timestamp": t + d})
            t += d + 10
    out = _bench().window_stats(str(p), real=2)
    assert out["window_rows"] == 4
    k = out["kernels"]["k_sweep<32, float>"]
    assert k["launches"] == 3 and k["total_ns"] == 26_000
    assert k["mean_us"] == pytest.approx(26_000 / 3 / 1e3)
    # the two longest launches are the real ones
    assert k["mean_us_of_longest"] == {"launches": 2, "mean_us": pytest.approx(12.0)}
    assert out["kernels"]["k_pairs"]["launches"] == 1
    # device time of the window: four kernels 10 ns apart
    assert out["busy_ms"] == pytest.approx(76_000 / 1e6)
    assert out["span_ms"] == pytest.approx(76_030 / 1e6)
    assert out["gaps_ms"]["under_10us"] == pytest.approx(30 / 1e6)
    assert out["gaps_ms"]["count_over_100us"] == 0


def test_counter_window(tmp_path):
    p = tmp_path / "counter_collection.csv"
    vals = [9, 9, 0, 0, 0, 300, 77, 1, 500, 0, 9, 9]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (name, v) in enumerate(zip(_names(), vals)):
            w.writerow({"Dispatch_Id": i + 1, "Kernel_Name": name, "Counter_Name": "FETCH_SIZE",
                        "Counter_Value": v})
    out = _bench().window_stats(str(p), real=2)
    c = out["kernels"]["k_sweep<32, float>"]["FETCH_SIZE"]
    assert c["dispatches"] == 3 and c["sum"] == 801
    assert c["mean"] == pytest.approx(267)
    assert c["per_real_launch"] == pytest.approx(400.5)
    assert "per_real_launch" not in out["kernels"]["k_pairs"]["FETCH_SIZE"]


def test_no_window_is_an_error(tmp_path):
    p = tmp_path / "kernel_trace.csv"
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerow({"Kernel_Name": "k_sweep<32, float>", "Start_Timestamp": 0, "End_Timestamp": 5})
    with pytest.raises(SystemExit):
        _bench().window_stats(str(p))
