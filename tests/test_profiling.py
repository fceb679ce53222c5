"""The rocprofv3 summary helpers (dawgs_amd/profiling.py) on small CSV files
written here: per-group means of a kernel's dispatches, and counter
scaling against a copy kernel of known size."""
import csv

import pytest

from dawgs_amd import profiling as P


def write_csv(path, header, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_trace_group_means(tmp_path):
    rows = []
    t = 0
    for i, dur in enumerate([10, 20, 30, 40, 100, 200, 300, 400]):
        rows.append(["k_a<double, 1>(int)", t, t + dur])
        rows.append(["other", t, t + 7])
        t += 1000
    write_csv(tmp_path / "x_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    assert P.trace_group_means(str(tmp_path), "k_a<double, 1>", 2, 2) == [35.0, 350.0]
    assert P.trace_group_means(str(tmp_path), "k_a<double, 1>", 4, 0) == [15.0, 35.0, 150.0, 350.0]
    with pytest.raises(ValueError):
        P.trace_group_means(str(tmp_path), "k_a<double, 1>", 3, 0)
    assert P.trace_tail_mean(str(tmp_path), "k_a", 4)["tail_mean_ns"] == 250.0


def test_counter_groups_and_traffic(tmp_path):
    f, w = tmp_path / "f", tmp_path / "w"
    f.mkdir()
    w.mkdir()
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    # a copy of 2 MiB moved: 1 MiB read (FETCH_SIZE counts half: 512 KiB), 1 MiB written
    write_csv(f / "p_counter_collection.csv", hdr,
              [[1, "neg_kernel", "FETCH_SIZE", 512], [2, "k_b", "FETCH_SIZE", 100], [3, "k_b", "FETCH_SIZE", 300]])
    write_csv(w / "p_counter_collection.csv", hdr,
              [[1, "neg_kernel", "WRITE_SIZE", 1024], [2, "k_b", "WRITE_SIZE", 50], [3, "k_b", "WRITE_SIZE", 70]])
    assert P.counter_group_means(str(f), "FETCH_SIZE", "k_b", 2, 0) == [100.0, 300.0]
    r = P.traffic(str(f), str(w), "k_b", "neg_kernel", 2 * 1024 * 1024, last=1)
    assert r["fetch_scale"] == 2.0 and r["write_scale"] == 1.0
    assert r["fetch_bytes"] == 300 * 1024 * 2 and r["write_bytes"] == 70 * 1024
