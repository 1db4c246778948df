"""The profile summaries' arithmetic (cfd__amd.rocprof_summary) on synthetic
rocprofv3 CSV files with known answers: the copy-kernel calibration of
FETCH_SIZE / WRITE_SIZE, the traffic ratio, the average over the last N
dispatches of the main instantiation and the fraction of the 8 TB/s peak."""
import csv
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
STEP = "void cfd::k_step<cfd::StoreF32, false, 512, 9>(args)"
STEP_M = "void cfd::k_step<cfd::StoreF32, true, 512, 9>(args)"
COPY = "cfd::k_copy(args)"


def _write(path, header, rows):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_summary_arithmetic(tmp_path):
    cells, bpc, copy_bytes = 1000000, 152, 1 << 30
    # trace: 10 dispatches of the step, the first 4 slow (200 us), the last 6 at 100 us
    durs = [200_000] * 4 + [100_000] * 6
    t, trace = 0, []
    for d in durs:
        trace.append([STEP, t, t + d])
        t += d + 1000
    trace.append([STEP_M, t, t + 300_000])
    _write(tmp_path / "trace/run_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp",
                                                     "End_Timestamp"], trace)
    _write(tmp_path / "trace/run_kernel_stats.csv",
           ["Name", "Calls", "AverageNs", "MinNs", "MaxNs", "Percentage"],
           [[STEP, 10, sum(durs) / 10, 100_000, 200_000, 80.0],
            [STEP_M, 1, 300_000, 300_000, 300_000, 15.0],
            [COPY, 4, 350_000, 340_000, 360_000, 5.0]])
    # counters: FETCH_SIZE reads half the copy's bytes (factor 2), WRITE_SIZE all of them;
    # the step reads 80 MB and writes 76 MB in counter units (calibrated: 160 + 76 MB)
    kib = copy_bytes / 1024
    _write(tmp_path / "fetch/run_counter_collection.csv",
           ["Kernel_Name", "Counter_Name", "Counter_Value"],
           [[COPY, "FETCH_SIZE", kib / 2], [STEP, "FETCH_SIZE", 80e6 / 1024],
            [STEP, "FETCH_SIZE", 80e6 / 1024], [STEP_M, "FETCH_SIZE", 1.0]])
    _write(tmp_path / "write/run_counter_collection.csv",
           ["Kernel_Name", "Counter_Name", "Counter_Value"],
           [[COPY, "WRITE_SIZE", kib], [STEP, "WRITE_SIZE", 76e6 / 1024],
            [STEP_M, "WRITE_SIZE", 1.0]])
    out = tmp_path / "s.md"
    subprocess.run([sys.executable, "-m", "cfd__amd.rocprof_summary", "--trace",
                    str(tmp_path / "trace"), "--fetch", str(tmp_path / "fetch"), "--write",
                    str(tmp_path / "write"), "--kernel", "k_step", "--cells", str(cells),
                    "--bytes-per-cell", str(bpc), "--copy-bytes", str(copy_bytes),
                    "--command", "python3 bench.py", "--out", str(out), "--last", "6"],
                   cwd=ROOT, check=True, capture_output=True, timeout=60)
    text = out.read_text()

    def val(key):
        return float(re.search(rf"- {key}: ([0-9.]+)", text).group(1))
    alg = bpc * cells
    assert val("algorithmic_bytes_per_launch") == alg
    assert abs(val("traffic_bytes_per_launch") - (2 * 80e6 + 76e6)) < 1.0
    assert abs(val("traffic / algorithmic") - (2 * 80e6 + 76e6) / alg) < 1e-4
    assert val("avg_kernel_us") == 100.0           # the last 6 dispatches only
    assert abs(val("fraction_of_8TBps_peak") - alg / 100e-6 / 8e12) < 1e-4
    assert "last 6 of 10 calls" in text
    assert "factor 2.0000" in text and "factor 1.0000" in text
