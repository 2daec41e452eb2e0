"""Host tests of bench.py's counter summary: the measured blocks of the counters
workload are found between the marker copies, scaled by the calibration
copies, and the bench line takes the traffic of the layout and block order
the autotuner chose."""
import csv
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _write_pass(directory, counter, scale, per_block):
    """A counter_collection.csv laid out as the counters workload launches its
    kernels: set-up kernels, then per block 3 warm steps, a marker (a fill and
    two copies), 10 steps (step kernel + a small side kernel), a closing
    marker; at the end a fill and four 1 GiB copies (the probe's warm-up copy
    and three timed ones).  Values are in KiB as rocprofv3 writes them, the
    counter reading `scale` times the bytes."""
    rows, n = [], [0]

    def add(name, nbytes):
        n[0] += 1
        rows.append({"Dispatch_Id": n[0], "Kernel_Name": name, "Counter_Name": counter,
                     "Counter_Value": nbytes * scale / 1024.0})

    add("void lbm3::k_init(args)", 5e6)
    for (st, key, remap), nbytes in zip(bench.counter_segments(), per_block):
        for _ in range(3):
            add("void lbm3::k_step_%s(args)" % key, 1.0)   # warm steps: must not count
        add("__amd_rocclr_fillBuffer", 4096)
        add("lbm3::k_copy(args)", 4096)
        add("lbm3::k_copy(args)", 4096)
        for _ in range(10):
            add("void lbm3::k_step_%s(args)" % key, nbytes)
            add("void lbm3::k_collide_wall(args)", 1e6)
        add("__amd_rocclr_fillBuffer", 4096)
        add("lbm3::k_copy(args)", 4096)
        add("lbm3::k_copy(args)", 4096)
    add("__amd_rocclr_fillBuffer", 1 << 30)
    for _ in range(4):
        add("lbm3::k_copy(args)", 1 << 30)
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "x_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)


def test_counter_summary_reads_each_block(tmp_path):
    order = bench.counter_segments()
    fetch = [4e8 + 1e6 * i for i in range(len(order))]
    write = [3e8 + 2e6 * i for i in range(len(order))]
    # FETCH_SIZE reads half the bytes (as on the MI355X), WRITE_SIZE all of them
    _write_pass(str(tmp_path / "f"), "FETCH_SIZE", 0.5, fetch)
    _write_pass(str(tmp_path / "w"), "WRITE_SIZE", 1.0, write)
    out = str(tmp_path / "t.json")
    bench.counters_summary(str(tmp_path / "f"), str(tmp_path / "w"), out)
    d = json.load(open(out))
    assert d["calibration"]["fetch_factor"] == 2.0 and d["calibration"]["write_factor"] == 1.0
    for i, (st, key, remap) in enumerate(order):
        b = d[key]["by_block_order"][str(remap)]
        assert b["fetch_bytes"] == round(fetch[i]) and b["write_bytes"] == round(write[i])
        assert d[key]["cells"] == bench.BULK_CELLS_SPHERE


def test_counter_summary_refuses_a_pass_of_another_workload(tmp_path):
    order = bench.counter_segments()
    _write_pass(str(tmp_path / "f"), "FETCH_SIZE", 0.5, [1e8] * len(order))
    _write_pass(str(tmp_path / "w"), "WRITE_SIZE", 1.0, [1e8] * (len(order) - 1))
    with pytest.raises(SystemExit):
        bench.counters_summary(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "t.json"))


def test_traffic_follows_the_chosen_layout_and_block_order():
    d = json.load(open(os.path.join(ROOT, "profiles", "hbm_traffic_sphere.json")))
    cells = bench.BULK_CELLS_SPHERE
    t = {"open_group": 1, "open_split": 0, "overlap": 3, "xcd_remap": 3}
    assert bench.traffic_from_profiles("f32", cells, t) == d["f32_in_bulk_walls"]["by_block_order"]["3"]["bytes_per_launch"]
    t = dict(t, overlap=1, xcd_remap=2)
    assert bench.traffic_from_profiles("f16", cells, t) == d["f16_in_bulk"]["by_block_order"]["2"]["bytes_per_launch"]
    t = dict(t, open_split=3)
    assert bench.traffic_from_profiles("bf16", cells, t) == d["bf16_in_bulk_rotated"]["by_block_order"]["2"]["bytes_per_launch"]
    # no counter pass for the grouped layout or for f16c: no traffic figure
    assert bench.traffic_from_profiles("f16", cells, dict(t, open_group=16)) is None
    assert bench.traffic_from_profiles("f16c", cells, t) is None
    # another cell split than the one measured
    assert bench.traffic_from_profiles("f32", cells - 1, t) is None
