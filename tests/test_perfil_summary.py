"""The around-the-step profile summary (cfd__amd.perfil --resumo) on synthetic
inputs with known answers: GB/s and the peak fraction per kernel from the
trace averages and the algorithmic bytes, the macro kernels told apart
(k_macros is a part of k_pull_macros' name), and the calibrated traffic."""
import csv
import json
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PULL = "void cfd::k_pull_macros<cfd::StoreF32, 9>(args)"
MAC = "void cfd::k_macros<cfd::StoreF32, 9>(args)"
MASS = "void cfd::k_mass<cfd::StoreF32, 9>(args)"
RASTER = "(anonymous namespace)::k_raster(args)"
COPY = "cfd::k_copy(args)"


def _csv(path, header, rows):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_perfil_summary_arithmetic(tmp_path):
    N = 1_000_000
    j = {"grid": [320, 160, 128], "cells": N, "triangles": 10, "precisao": "fp32",
         "repeticoes": 3,
         "voxelizar": {"ms_raster": 0.5, "ms_flood": 0.4, "ms_total": 1.0, "flood_sweeps": 5.0,
                       "pairs": 1_000_000, "pairs_per_s": 2e9, "flood_bytes": 5_000_000,
                       "flood_GBps": 12.5},
         "forcas": {"us_per_call": 20.0, "link_cells": 10, "links": 30, "bytes": 200,
                    "GBps": 0.01},
         "massa": {"us_per_call": 90.0, "bytes": 77_000_000, "GBps": 855.0},
         "parede": {"us_per_call": 200.0, "bytes": 8 * N}}
    (tmp_path / "p.json").write_text(json.dumps(j))
    # trace averages: pull-macros 92 MB in 46 us (2000 GB/s), macros in 23 us, mass 77 MB in 77 us
    _csv(tmp_path / "trace/run_kernel_stats.csv",
         ["Name", "Calls", "AverageNs", "MinNs", "MaxNs", "Percentage"],
         [[PULL, 3, 46_000, 0, 0, 10], [MAC, 3, 23_000, 0, 0, 5], [MASS, 3, 77_000, 0, 0, 5],
          [RASTER, 3, 500_000, 0, 0, 50], [COPY, 4, 350_000, 0, 0, 30]])
    kib = (1 << 30) / 1024
    _csv(tmp_path / "fetch/run_counter_collection.csv",
         ["Kernel_Name", "Counter_Name", "Counter_Value"],
         [[COPY, "FETCH_SIZE", kib / 2], [PULL, "FETCH_SIZE", 40e6 / 1024],
          [MASS, "FETCH_SIZE", 38.5e6 / 1024]])
    _csv(tmp_path / "write/run_counter_collection.csv",
         ["Kernel_Name", "Counter_Name", "Counter_Value"],
         [[COPY, "WRITE_SIZE", kib], [PULL, "WRITE_SIZE", 16e6 / 1024],
          [MASS, "WRITE_SIZE", 0.0]])
    out = tmp_path / "s.md"
    subprocess.run([sys.executable, "-m", "cfd__amd.perfil", "--resumo", str(tmp_path / "trace"),
                    "--json", str(tmp_path / "p.json"), "--fetch", str(tmp_path / "fetch"),
                    "--write", str(tmp_path / "write"), "--comando", "python3 -m cfd__amd.perfil",
                    "--saida", str(out)], cwd=ROOT, check=True, capture_output=True, timeout=60)
    text = out.read_text()

    def row(name):
        line = next(l for l in text.splitlines() if l.startswith(f"| `{name}`"))
        return [c.strip() for c in line.strip("|").split("|")]
    pull, mac, mass = row(PULL), row(MAC), row(MASS)
    assert pull[3] == f"{92 * N:,}" and mac[3] == f"{92 * N:,}"
    assert pull[4] == "2000" and pull[5] == "0.250"
    assert mac[4] == "4000" and mac[5] == "0.500"
    assert mass[3] == "77,000,000" and mass[4] == "1000" and mass[5] == "0.125"
    # traffic: FETCH calibrated by the copy's factor 2, WRITE by 1
    m = re.search(r"- k_pull_macros: read ([0-9.]+) MB \+ write ([0-9.]+) MB", text)
    assert m and float(m.group(1)) == 80.0 and float(m.group(2)) == 16.0
    assert "rocprofv3 --pmc FETCH_SIZE" in text
