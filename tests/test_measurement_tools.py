"""CPU tests of the measurement tooling: bench.py's one-GPU contract and the
rocprof summaries (ttc_amd/profiling.py) on synthetic rocprofv3 CSV files."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from ttc_amd import profiling  # noqa: E402

K = "void ttc::transpose_rc_kernel<ttc::OpDD, 1, 1, true, false>(ttc::RCParams, double const*, double*, double, double)"
COPY = "__amd_rocclr_copyBuffer"


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_short_name_and_last_dispatches(tmp_path):
    assert profiling.short_name(K) == "ttc::transpose_rc_kernel<ttc::OpDD, 1, 1, true, false>"
    d = str(tmp_path / "trace")
    rows = []
    for i in range(10):  # launches 100 us long, 2 us apart; the first five are 50 us
        start = i * 102_000
        rows.append(["KERNEL_DISPATCH", i + 1, K, start, start + (50_000 if i < 5 else 100_000)])
    _write(os.path.join(d, "run_kernel_trace.csv"), ["Kind", "Dispatch_Id", "Kernel_Name", "Start_Timestamp",
                                                      "End_Timestamp"], rows)
    avg, lo, hi, span, n = profiling.last_dispatches(d, profiling.short_name(K), 4)
    assert (avg, lo, hi, n) == (100_000, 100_000, 100_000, 4)
    assert span == (9 * 102_000 + 100_000 - 6 * 102_000) / 4


def test_traffic_scales_counters_by_the_copy(tmp_path):
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    copy_bytes = 200 << 20
    # the copy reads half its bytes in FETCH_SIZE (128-byte requests counted as 64) and all in WRITE_SIZE
    _write(str(tmp_path / "cf" / "run_counter_collection.csv"), hdr,
           [[1, COPY, "FETCH_SIZE", copy_bytes / 2048.0]])
    _write(str(tmp_path / "cw" / "run_counter_collection.csv"), hdr, [[1, COPY, "WRITE_SIZE", copy_bytes / 1024.0]])
    # a kernel whose dispatches read 300 and 400 "counter" KiB and write 500 KiB; the last one counts with last=1
    _write(str(tmp_path / "f" / "run_counter_collection.csv"), hdr,
           [[1, K, "FETCH_SIZE", 300.0], [2, K, "FETCH_SIZE", 400.0]])
    _write(str(tmp_path / "w" / "run_counter_collection.csv"), hdr,
           [[1, K, "WRITE_SIZE", 500.0], [2, K, "WRITE_SIZE", 500.0]])
    res = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                            copy_bytes)
    assert res["calibration"]["fetch_scale"] == 2.0 and res["calibration"]["write_scale"] == 1.0
    k = res["kernels"][profiling.short_name(K)]
    assert k["launches"] == 2 and k["fetch_bytes"] == round(350 * 1024 * 2) and k["write_bytes"] == 500 * 1024
    last = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                             copy_bytes, last=1)["kernels"][profiling.short_name(K)]
    assert last["fetch_bytes"] == 400 * 1024 * 2
    json.dumps(res)
