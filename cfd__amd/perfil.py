"""Timing of the kernels around the step on the 320x160x128 vehicle: the
voxelizer (voxelizar, src/geom/voxel.js:376-433), the momentum-exchange
force reduction (solver.js:454-479) and the mass check.

Each piece is timed on the device (the voxelizer's own events) or by the
host around a synchronous call (forces, mass: launch, two-stage float64
reduction and the 24- or 16-byte readback), and reported with the bytes or
work its algorithm needs, so that it can be set against the HBM peak:

* rasterization: (triangle, cell) pairs tested, the cells of each triangle's
  box padded by the band (voxel.js:195-201), ~100 float64 FLOP each;
* flood (seal dilation, seed, tile sweeps, skin, final): bytes of the cell
  arrays each kernel must read and write once (1-byte flags, 4-byte dist2);
* forces: the body-link list (8 bytes per link cell: the cell and its mask
  of solid directions, built when the types are set) and one population per
  link;
* mass: 19 populations and the type of every cell;
* wall distance of the learned closure (cfd_compute_wall_distance): per
  cell the type and the row distance (rows), the row distance and the plane
  distance (columns), the plane distance and the result (planes).

usage: python -m cfd__amd.perfil [--repeticoes R] [--precisao fp32|fp16] [--copia]
Prints one JSON object.  Run it under rocprofv3 --kernel-trace --stats for
per-kernel averages, then
  python -m cfd__amd.perfil --resumo TRACE --json OUT.json [--fetch D --write D]
      --comando "PROGRAM" --saida profiles/x.md
writes the markdown summary (kernel table, algorithmic bytes per call and
GB/s from the kernel-trace averages, HBM traffic from the counters).  PROGRAM
is the profiled command (python3 -m cfd__amd.perfil --repeticoes 20 --copia);
the summary lists every command of the recipe.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

GRID = (320, 160, 128)


def pares_raster(pos: np.ndarray, idx: np.ndarray, grade, banda: float) -> int:
    """Cells tested by the rasterizer: per triangle the box of its vertices
    padded by the band, floor/ceil, clamped to the grid (voxel.js:195-201)."""
    p = np.asarray(pos, dtype=np.float32).reshape(-1, 3).astype(np.float64)
    t = np.asarray(idx, dtype=np.int64).reshape(-1, 3)
    v = p[t]                                        # [nt, 3 vertices, 3 axes]
    n = np.array(grade, dtype=np.float64)
    lo = np.maximum(np.floor(v.min(axis=1) - banda), 0.0)
    hi = np.minimum(np.ceil(v.max(axis=1) + banda), n - 1)
    ext = np.clip(hi - lo + 1, 0, None)
    return int(np.prod(ext, axis=1).sum())


def contar_links(tipo: np.ndarray, grade) -> tuple[int, int]:
    """(cells flagged BODY_LINK, links): FLUIDO cells with SOLIDO neighbours
    over the 18 moving directions, periodic wrap as the kernels use."""
    from .lattice import C
    nx, ny, nz = grade
    t = np.asarray(tipo).reshape(nz, ny, nx)
    fluido = t == 0
    tem = np.zeros_like(fluido)
    links = 0
    for i in range(1, 19):
        c = C[i]
        nb = np.roll(t, shift=(-c[2], -c[1], -c[0]), axis=(0, 1, 2)) == 1
        tem |= fluido & nb
        links += int(np.count_nonzero(fluido & nb))
    return int(np.count_nonzero(tem)), links


def _stats(d):
    import csv
    from pathlib import Path
    rows = list(csv.DictReader(open(Path(d) / "run_kernel_stats.csv")))
    return {r["Name"]: (int(r["Calls"]), float(r["AverageNs"]) / 1e3) for r in rows}


def _counter(d, name):
    import csv
    from collections import defaultdict
    from pathlib import Path
    agg = defaultdict(list)
    for r in csv.DictReader(open(Path(d) / "run_counter_collection.csv")):
        if r["Counter_Name"] == name:
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def resumo(a) -> None:
    """Markdown summary of a profiled run of this module."""
    from pathlib import Path
    j = json.loads(Path(a.json).read_text().strip().splitlines()[-1])
    st = _stats(a.resumo)
    N = j["cells"]

    def avg(key):
        for k, (n, us) in st.items():
            if key in k:
                return n, us
        return 0, float("nan")

    # algorithmic bytes per launch of each cell-array kernel (1-byte arrays)
    per_cell = {"k_dilate": 2, "k_seed": 2, "k_flood_tile": 3, "k_skin": 3, "k_final": 3,
                "k_prep_types": 4 * 19 * 2 + 1, "k_wd_rows": 2, "k_wd_cols": 3,
                "k_wd_planes": 3}
    # the macro kernels: 19 populations read, one float4 written per cell
    # (the type byte excluded, as for the step); k_pull_macros first, since
    # k_macros is a part of its name
    bpp = 4 if j["precisao"] == "fp32" else 2
    per_cell = {"k_pull_macros": 19 * bpp + 16, "k_macros": 19 * bpp + 16, **per_cell}
    lines = [f"# Kernels around the step, {N // (160 * 128)}x160x128 vehicle "
             f"({N / 1e6:.2f} M cells, {j['triangles']} triangles, {j['precisao']})", "",
             "Commands (run from the repository root on one MI355X):", "", "```",
             "python3 -m cfd__amd.perfil > perfil.json",
             f"rocprofv3 --kernel-trace --stats -d <trace> -o run --output-format csv -- {a.comando}",
             f"rocprofv3 --pmc FETCH_SIZE -d <fetch> -o run --output-format csv -- {a.comando}",
             f"rocprofv3 --pmc WRITE_SIZE -d <write> -o run --output-format csv -- {a.comando}",
             "python -m cfd__amd.perfil --resumo <trace> --json perfil.json --fetch <fetch> "
             f"--write <write> --comando '{a.comando}' --saida {a.saida}",
             "```", "", "## Kernel trace (--stats)", "",
             "| kernel | calls | avg us | algorithmic bytes / call | GB/s | of 8 TB/s |",
             "|---|---|---|---|---|---|"]
    fo, ma = j["forcas"], j["massa"]
    alg = {"k_forces": fo["bytes"], "k_mass": ma["bytes"]}
    for k, v in per_cell.items():
        alg[k] = v * N
    for name, (n, us) in sorted(st.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        short = next((k for k in alg if k in name), None)
        if short and short != "k_prep_types":
            b = alg[short]
            gbs = b / (us * 1e-6) / 1e9
            lines.append(f"| `{name}` | {n} | {us:.2f} | {b:,} | {gbs:.0f} | {gbs / 8000:.3f} |")
        else:
            lines.append(f"| `{name}` | {n} | {us:.2f} | | | |")
    vr = j["voxelizar"]
    n_r, us_r = avg("k_raster")
    lines += ["", "## Per piece", "",
              f"- rasterization (`k_raster`): {vr['pairs']:,} (triangle, cell) pairs per call, "
              f"{us_r:.1f} us in the trace -> {vr['pairs'] / (us_r * 1e-6) / 1e9:.2f} G pairs/s "
              "(float64 distance + 13-axis overlap per pair: bound by float64 ALU, not HBM)",
              f"- flood side: {vr['flood_sweeps']:.0f} tile sweeps, {vr['flood_bytes']:,} "
              f"algorithmic bytes, {vr['ms_flood'] * 1e3:.0f} us device time (events) -> "
              f"{vr['flood_GBps']:.0f} GB/s; whole voxelization {vr['ms_total']:.3f} ms",
              f"- forces: {fo['link_cells']:,} link cells, {fo['links']:,} links; host-timed "
              f"call (launch, two-stage reduction, readback) {fo['us_per_call']:.1f} us",
              f"- mass: host-timed call {ma['us_per_call']:.1f} us, "
              f"{ma['bytes']:,} algorithmic bytes"]
    if "parede" in j:
        pa = j["parede"]
        lines.append(f"- wall distance: host-timed call (scratch allocation, table upload, "
                     f"three passes, free) {pa['us_per_call']:.1f} us; {pa['bytes']:,} "
                     "algorithmic bytes over the three passes")
    if a.fetch and a.write:
        fe, wr = _counter(a.fetch, "FETCH_SIZE"), _counter(a.write, "WRITE_SIZE")
        kc = next(k for k in fe if "k_copy" in k)
        f_fac = (1 << 30) / (fe[kc] * 1024)
        w_fac = (1 << 30) / (wr[next(k for k in wr if "k_copy" in k)] * 1024)
        lines += ["", "## HBM traffic (--pmc, per dispatch)", "",
                  f"- copy-kernel calibration (1 GiB each way): FETCH factor {f_fac:.4f}, "
                  f"WRITE factor {w_fac:.4f}", ""]
        for short in ("k_forces", "k_mass", "k_pull_macros", "k_flood_tile", "k_raster"):
            kf = next((k for k in fe if short in k), None)
            kw = next((k for k in wr if short in k), None)
            if kf is None or kw is None:
                continue
            rd, wb = fe[kf] * 1024 * f_fac, wr[kw] * 1024 * w_fac
            extra = f" (algorithmic {alg[short]:,})" if short in alg else ""
            lines.append(f"- {short}: read {rd / 1e6:.2f} MB + write {wb / 1e6:.2f} MB{extra}")
    Path(a.saida).write_text("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--repeticoes", type=int, default=50)
    ap.add_argument("--precisao", choices=("fp32", "fp16"), default="fp32")
    ap.add_argument("--copia", action="store_true",
                    help="also run the 1 GiB copy kernel (calibrates rocprofv3 --pmc byte counters)")
    ap.add_argument("--resumo", default=None, help="kernel-trace directory to summarize")
    ap.add_argument("--json", default=None)
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--comando", default="")
    ap.add_argument("--saida", default=None)
    a = ap.parse_args()
    if a.resumo:
        return resumo(a)
    if a.copia:
        from ._lib import copy_bandwidth
        copy_bandwidth(1 << 30, 3)

    from . import meshes, tunel
    from .solver import SolverMI355X
    from .voxel import voxelizar

    nx, ny, nz = GRID
    g = {"nx": nx, "ny": ny, "nz": nz}
    N = nx * ny * nz
    pos, idx, info = meshes.vehicle(seed=7)
    banda = 1.5
    voxelizar(pos, idx, g, banda=banda, selarFrestas=True, comSdf=False)   # warm-up
    ras, fl, tot, its = [], [], [], []
    for _ in range(a.repeticoes):
        v = voxelizar(pos, idx, g, banda=banda, selarFrestas=True, comSdf=False)
        st = v["stats"]
        ras.append(st["msRasterizacao"])
        fl.append(st["msInundacao"])
        tot.append(st["msTotal"])
        its.append(st["iteracoesInundacao"])
    pares = pares_raster(pos, idx, GRID, banda)
    # flood-side bytes per voxelization: three dilation passes (read one
    # array, write one), seed (read barrier, write ext), each tile sweep (read
    # barrier + ext, write ext), skin (read before + flags, write after), final
    # (read flags + ext, write tipo); no distance field without the sdf
    sweeps = float(np.median(its))
    bytes_flood = N * (3 * 2 + 2 + 3 * sweeps + 3 + 3)

    tipo, esponja, _ = tunel.montarTipos(g, v["tipo"], esteira=False)
    link_cells, links = contar_links(tipo, GRID)
    s = SolverMI355X.criar(g, precisao=a.precisao)
    try:
        s.definirTipos(tipo)
        s.configurar(omegaPlus=1.9, lesCs=0.1, inletU=[0.08, 0.0, 0.0], esponja=esponja)
        s.inicializar()
        s.passo(100)
        s.sincronizar()
        s.forcaBruta()
        t0 = time.perf_counter()
        for _ in range(a.repeticoes):
            s.forcaBruta()
        t_forca = (time.perf_counter() - t0) / a.repeticoes
        s.massa()
        t0 = time.perf_counter()
        for _ in range(a.repeticoes):
            s.massa()
        t_massa = (time.perf_counter() - t0) / a.repeticoes
        s.calcularDistanciaParede()
        t0 = time.perf_counter()
        for _ in range(a.repeticoes):
            s.calcularDistanciaParede()
        t_parede = (time.perf_counter() - t0) / a.repeticoes
        # the macro field: read after each step (k_pull_macros: that step's
        # pull and moments over the set it read), and atualizarMacros
        # (k_macros: the moments of the current set); kernel times come from
        # the trace
        for _ in range(a.repeticoes):
            s.passo(1)
            s.lerMacros()
        for _ in range(a.repeticoes):
            s.atualizarMacros()
        s.sincronizar()
    finally:
        s.destruir()
    bpp = 4 if a.precisao == "fp32" else 2
    bytes_forca = 8 * link_cells + bpp * links
    non_wall = int(np.count_nonzero(~np.isin(tipo, (1, 4, 7))))
    bytes_massa = N + 19 * bpp * non_wall
    print(json.dumps({
        "grid": list(GRID), "cells": N, "triangles": int(idx.size // 3), "precisao": a.precisao,
        "repeticoes": a.repeticoes,
        "voxelizar": {
            "ms_raster": float(np.median(ras)), "ms_flood": float(np.median(fl)),
            "ms_total": float(np.median(tot)), "flood_sweeps": sweeps,
            "pairs": pares, "pairs_per_s": pares / (np.median(ras) * 1e-3),
            "flood_bytes": bytes_flood,
            "flood_GBps": bytes_flood / (np.median(fl) * 1e-3) / 1e9,
        },
        "forcas": {"us_per_call": t_forca * 1e6, "link_cells": link_cells, "links": links,
                   "bytes": bytes_forca, "GBps": bytes_forca / t_forca / 1e9},
        "massa": {"us_per_call": t_massa * 1e6, "bytes": bytes_massa,
                  "GBps": bytes_massa / t_massa / 1e9},
        "parede": {"us_per_call": t_parede * 1e6, "bytes": 8 * N},
    }))


if __name__ == "__main__":
    main()
