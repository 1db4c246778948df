"""Summarize rocprofv3 CSV output into the short markdown kept under profiles/.

Inputs are the directories written by three runs of the same command:
  rocprofv3 --kernel-trace --stats -d <trace> -o run --output-format csv -- <cmd>
  rocprofv3 --pmc FETCH_SIZE      -d <fetch> -o run --output-format csv -- <cmd>
  rocprofv3 --pmc WRITE_SIZE      -d <write> -o run --output-format csv -- <cmd>

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are calibrated against the copy
kernel k_copy that the same command runs on a known byte count
(cfd_copy_bandwidth, read = write = copy_bytes per dispatch): the factor is
known bytes / counter, applied to the step kernel's counters.

usage: python -m cfd__amd.rocprof_summary --trace D --fetch D --write D
           --kernel k_step --cells N --bytes-per-cell 152 --copy-bytes 1073741824
           --command "..." --out profiles/x.md [--last N]

--last N takes the roofline's average duration from the last N dispatches of
the main instantiation in the per-dispatch trace (a bench's timed region after
its warm-up) instead of the average over all of them.

--dram D adds a fourth run,
  rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B TCC_EA0_WRREQ_WRITE_DRAM_32B -d <dram> ...
whose counters count DRAM requests in 32-byte units whatever the request size
(a 128-byte request counts 4), so they give bytes without calibration.  On
gfx950, FETCH_SIZE's 128-byte term (TCC_BUBBLE) reads zero and every
128-byte read counts as 64 bytes, which is why the copy-kernel factor is 2;
the exact counters confirm the calibrated figures.
"""
from __future__ import annotations

import argparse
import csv
from collections import defaultdict
from pathlib import Path


def _stats(trace_dir: Path):
    rows = list(csv.DictReader(open(trace_dir / "run_kernel_stats.csv")))
    return [(r["Name"], int(r["Calls"]), float(r["AverageNs"]), float(r["MinNs"]),
             float(r["MaxNs"]), float(r["Percentage"])) for r in rows]


def _last_avg_ns(trace_dir: Path, name: str, n: int):
    """Mean duration (ns) of the last n dispatches of kernel `name`."""
    rows = [r for r in csv.DictReader(open(trace_dir / "run_kernel_trace.csv"))
            if r["Kernel_Name"] == name]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    d = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows[-n:]]
    return sum(d) / len(d), len(d)


def _counter(d: Path, name: str):
    agg = defaultdict(list)
    for r in csv.DictReader(open(d / "run_counter_collection.csv")):
        if r["Counter_Name"] == name:
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def _pick(d: dict, key: str):
    for k, v in d.items():
        if key in k:
            return k, v
    raise KeyError(key)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", type=Path, required=True)
    ap.add_argument("--fetch", type=Path, required=True)
    ap.add_argument("--write", type=Path, required=True)
    ap.add_argument("--kernel", default="k_step")
    ap.add_argument("--cells", type=int, required=True)
    ap.add_argument("--steps", type=int, default=1,
                    help="steps per launch (a batch launch runs many)")
    ap.add_argument("--bytes-per-cell", type=int, default=152)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--command", required=True)
    ap.add_argument("--title", default="")
    ap.add_argument("--out", type=Path, required=True)
    ap.add_argument("--last", type=int, default=0,
                    help="average only the last N dispatches of the main instantiation")
    ap.add_argument("--dram", type=Path, default=None,
                    help="directory of a TCC_EA0_RDREQ_DRAM_32B / TCC_EA0_WRREQ_WRITE_DRAM_32B run")
    ap.add_argument("--match", default="",
                    help="the main instantiation's name must also contain this text "
                         "(e.g. 'StoreF32, false, 512')")
    a = ap.parse_args()

    st = _stats(a.trace)
    fetch, write = _counter(a.fetch, "FETCH_SIZE"), _counter(a.write, "WRITE_SIZE")
    _, cf = _pick(fetch, "k_copy")
    _, cw = _pick(write, "k_copy")
    kib = a.copy_bytes / 1024
    ff, fw = kib / cf, kib / cw
    # the step kernel without the macro store is the one timed per step
    step_rows = [r for r in st if a.kernel in r[0] and a.match in r[0]]
    main_row = max(step_rows, key=lambda r: r[1])
    kname = main_row[0]
    # the counter entry of the same instantiation as the trace's main row
    # (the macro-writing variant has `true` in its template arguments)
    def same(dct):
        best = None
        for k, v in dct.items():
            if k == kname or (best is None and a.kernel in k and a.match in k and
                              ("true" in k) == ("true" in kname)):
                best = v
        return best
    fk, wk = same(fetch), same(write)
    traffic = (fk * ff + wk * fw) * 1024
    alg = a.bytes_per_cell * a.cells * a.steps
    avg_s = main_row[2] * 1e-9
    avg_note = f"main instantiation, {main_row[1]} calls"
    if a.last > 0:
        avg_ns, n = _last_avg_ns(a.trace, kname, a.last)
        avg_s = avg_ns * 1e-9
        avg_note = f"main instantiation, last {n} of {main_row[1]} calls"
    lines = [
        f"# {a.title or a.kernel}",
        "",
        "Commands (run from the repository root on one MI355X):",
        "",
        "```",
        f"rocprofv3 --kernel-trace --stats -d <trace> -o run --output-format csv -- {a.command}",
        f"rocprofv3 --pmc FETCH_SIZE -d <fetch> -o run --output-format csv -- {a.command}",
        f"rocprofv3 --pmc WRITE_SIZE -d <write> -o run --output-format csv -- {a.command}",
        *([f"rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B TCC_EA0_WRREQ_WRITE_DRAM_32B -d <dram> "
           f"-o run --output-format csv -- {a.command}"] if a.dram is not None else []),
        f"python -m cfd__amd.rocprof_summary --trace <trace> --fetch <fetch> --write <write> "
        f"--kernel {a.kernel} --cells {a.cells} --steps {a.steps} "
        f"--bytes-per-cell {a.bytes_per_cell} "
        f"--copy-bytes {a.copy_bytes} --command '...' --out {a.out}"
        + (f" --last {a.last}" if a.last > 0 else "")
        + (f" --match '{a.match}'" if a.match else "")
        + (" --dram <dram>" if a.dram is not None else ""),
        "```",
        "",
        "## Kernel trace (--stats)",
        "",
        "| kernel | calls | avg us | min us | max us | % time |",
        "|---|---|---|---|---|---|",
    ]
    for name, calls, avg, mn, mx, pct in sorted(st, key=lambda r: -r[5])[:12]:
        short = name.split("(")[0][:90]
        lines.append(f"| `{short}` | {calls} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | "
                     f"{mx / 1e3:.2f} | {pct:.2f} |")
    lines += [
        "",
        "## HBM traffic (--pmc, per dispatch)",
        "",
        f"- copy-kernel calibration: k_copy moves {a.copy_bytes} B each way; FETCH_SIZE reads "
        f"{cf:.1f} KiB -> factor {ff:.4f}; WRITE_SIZE reads {cw:.1f} KiB -> factor {fw:.4f}",
        f"- {a.kernel}: FETCH_SIZE {fk:.1f} KiB, WRITE_SIZE {wk:.1f} KiB (raw)",
        f"- {a.kernel}: calibrated read {fk * ff * 1024 / 1e6:.1f} MB + write "
        f"{wk * fw * 1024 / 1e6:.1f} MB per launch",
        "",
    ]
    if a.dram is not None:
        dr, dw = _counter(a.dram, "TCC_EA0_RDREQ_DRAM_32B"), _counter(a.dram,
                                                                  "TCC_EA0_WRREQ_WRITE_DRAM_32B")
        _, crd = _pick(dr, "k_copy")
        _, cwr = _pick(dw, "k_copy")
        rk, wk2 = same(dr), same(dw)
        lines += [
            f"- exact DRAM requests (TCC_EA0_RDREQ_DRAM_32B / TCC_EA0_WRREQ_WRITE_DRAM_32B, "
            f"32-byte units, no calibration): k_copy read {crd * 32 / a.copy_bytes:.4f} and "
            f"write {cwr * 32 / a.copy_bytes:.4f} of its bytes; {a.kernel} read "
            f"{rk * 32 / 1e6:.1f} MB + write {wk2 * 32 / 1e6:.1f} MB per launch = "
            f"{(rk + wk2) * 32 / (a.bytes_per_cell * a.cells * a.steps):.4f} of algorithmic "
            f"(read {rk * 32 / (a.bytes_per_cell / 2 * a.cells * a.steps):.4f}, write "
            f"{wk2 * 32 / (a.bytes_per_cell / 2 * a.cells * a.steps):.4f})",
            "",
        ]
    else:
        lines += [""]
    lines += [
        "## Roofline",
        "",
        f"- algorithmic_bytes_per_launch: {alg}  ({a.bytes_per_cell} B/cell x {a.cells} cells"
        f" x {a.steps} steps)",
        f"- traffic_bytes_per_launch: {traffic:.0f}",
        f"- traffic / algorithmic: {traffic / alg:.4f}",
        f"- avg_kernel_us: {avg_s * 1e6:.2f}  ({avg_note})",
        f"- achieved_algorithmic_GBps: {alg / avg_s / 1e9:.1f}",
        f"- fraction_of_8TBps_peak: {alg / avg_s / 8e12:.4f}",
        f"- cell updates per second: {a.cells * a.steps / avg_s / 1e6:.0f} MLUPS",
        "",
    ]
    a.out.parent.mkdir(parents=True, exist_ok=True)
    a.out.write_text("\n".join(lines))
    print(a.out.read_text())


if __name__ == "__main__":
    main()
