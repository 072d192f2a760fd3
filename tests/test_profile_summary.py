"""The rocprof summariser behind profiles/: calibration of FETCH/WRITE by
the copy of known size, the main-kernel pick, and the counter mode's
derived figures, on small hand-made CSVs whose answers are known."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "da-spmmul-kernel_amd", "profile_summary.py")


def write_counters(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writerows(rows)


def run(*args):
    out = subprocess.run([sys.executable, TOOL, *args], capture_output=True, text=True, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_fetch_write_calibration_and_main_kernel(tmp_path):
    copy = "__amd_rocclr_copyBuffer"
    kern = "void dspmm::k_nnz_tiled<float, 4, 16, 4>(dspmm::ItemDesc, int)"
    # the copy of 1 MiB reads 2048 KiB by the counter (x0.5) and writes 1024 KiB (x1)
    write_counters(tmp_path / "fetch.csv", [(copy, "FETCH_SIZE", 2048), (copy, "FETCH_SIZE", 8),
                                            (kern, "FETCH_SIZE", 400), (kern, "FETCH_SIZE", 600)])
    write_counters(tmp_path / "write.csv", [(copy, "WRITE_SIZE", 1024), (kern, "WRITE_SIZE", 100)])
    with open(tmp_path / "stats.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "AverageNs"])
        w.writerow([copy, 2, 1000.0])
        w.writerow([kern, 2, 5000.0])
    meta = {"workload": "c2", "n": 64, "bf16": False, "nnz": 10,
            "config": {"variant": "nnz_tiled", "vec": 4, "lanes": 16, "tile_nnz": 64, "unroll": 4},
            "copy_bytes": 1 << 20, "algorithmic_bytes": 1 << 20}
    (tmp_path / "meta.log").write_text("noise\n" + json.dumps(meta) + "\n")
    res_path = tmp_path / "s.json"
    run("--trace", str(tmp_path / "stats.csv"), "--fetch", str(tmp_path / "fetch.csv"),
        "--write", str(tmp_path / "write.csv"), "--meta", str(tmp_path / "meta.log"),
        "--command", "cmd", "--out", str(res_path))
    res = json.loads(res_path.read_text())
    assert res["calibration"]["fetch"] == 0.5 and res["calibration"]["write"] == 1.0
    assert res["main_kernel"] == "k_nnz_tiled<float, 4, 16, 4>"
    k = res["kernels"]["k_nnz_tiled<float, 4, 16, 4>"]
    assert k["FETCH_SIZE_KiB"] == 500                       # mean over the dispatches
    assert k["traffic_bytes"] == 500 * 1024 * 0.5 + 100 * 1024 * 1.0


def test_counter_mode_derived_figures(tmp_path):
    kern = "void dspmm::k_row_seq<float, 4, 4>(dspmm::RowSeqLaunch, int)"
    cus, xcds = 4, 2
    gui = 2000.0                                           # 1000 cycles per XCD
    rows = [("k_fill_zero_u32(unsigned int*, long)", "GRBM_GUI_ACTIVE", 5.0),
            (kern, "GRBM_GUI_ACTIVE", gui),
            (kern, "TA_TA_BUSY_sum", 0.5 * 1000 * cus),
            (kern, "TD_TD_BUSY_sum", 0.25 * 1000 * cus),
            (kern, "TCP_TOTAL_CACHE_ACCESSES_sum", 1000.0),
            (kern, "TCP_TCC_READ_REQ_sum", 400.0),
            (kern, "TCP_TCC_READ_REQ_LATENCY_sum", 400.0 * 300),
            (kern, "TCC_HIT_sum", 300.0), (kern, "TCC_MISS_sum", 100.0),
            (kern, "SQ_WAVES", 10.0), (kern, "SQ_INSTS_VMEM_RD", 250.0)]
    write_counters(tmp_path / "a_p1_counter_collection.csv", rows)
    out = tmp_path / "c.md"
    res = run("--pmc", f"cfg={tmp_path}/a_p*_counter_collection.csv", "--cus", str(cus),
              "--xcds", str(xcds), "--title", "t", "--command", "c", "--out", str(out))
    assert res["kernels"] == ["k_row_seq<float, 4, 4>"]     # fills are left out
    text = out.read_text()
    want = {"TA busy": "0.5", "TD busy": "0.25", "L1 hit rate": "0.6", "L2 hit rate": "0.75",
            "L1-miss latency (cycles)": "300",
            # 400 misses / 4 CUs / 1000 cycles x 300 cycles = 30 in flight per CU
            "L1 misses in flight per CU": "30", "vector-memory reads per wave": "25"}
    for name, value in want.items():
        assert f"| {name} | {value} |" in text, name


def test_committed_headline_profile_feeds_the_bench():
    """bench.py reports roofline.traffic only when the measured configuration
    equals the profiled one; the committed headline profile's configuration
    must compare equal field by field to a configuration of the current
    struct (a field added later takes its zero default), and its traffic
    must be readable for exactly that configuration."""
    sys.path.insert(0, ROOT)
    import bench
    import dspmm as D
    prof = bench.profiled_config(bench.HEADLINE_TAG)
    assert prof is not None
    assert prof == D.config_from_dict(prof).as_dict()
    assert set(prof) == set(D.make_config().as_dict())
    traffic = bench.traffic_from_profile(bench.HEADLINE_TAG, prof)
    assert isinstance(traffic, int) and traffic > 0
    other = dict(prof, unroll=prof["unroll"] + 1)
    assert bench.traffic_from_profile(bench.HEADLINE_TAG, other) is None
