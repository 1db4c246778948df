"""Benchmark: D3Q19 TRT-LES on one MI355X, synthetic vehicle at 320x160x128.

Workload (the large plain-method case): FP32 populations, 320x160x128 cells
(6 553 600), u_in = 0.08, seeded superellipsoid vehicle with four wheels
(200 000 triangles, voxelized on the GPU) on a no-slip floor (PAREDE),
free-slip sides and top, Smagorinsky Cs = 0.1, Lambda = 3/16.  The requested
Re ~ 1e5 needs omega = 1.9991; the solver clamps omega at 1.90 as the
reference's Units does (units.js:157-164), so the run is at Re_lattice ~ 876.

A "step" is one lattice update of the whole domain.  The timed region holds K
steps in chunks of 100 (the smoothstep inflow ramp reconfigures between
chunks, as the reference's passo() does) plus a momentum-exchange force sample
after every chunk of the last 5000 steps; the line reports the sample window
and flags a window that starts inside the 6400-step inflow ramp.  value =
MLUPS = cells * K / wall.  The step kernel runs in the launch shape the
library picks for this population set; a warm-up long enough for the
launch-order autotuner (3000 steps) times the candidate shapes on live steps
first and keeps the fastest (results are bit-identical in every shape), and
--ajuste fixes one.

After the timed region, and outside it: the precision check (the validation
cases of the reference's validacao.html, sphere and cylinder at Re 100, plus
the 160x80x64 sphere, each in float32, float16 and float16 with stochastic
rounding: Cd and St error of the float16 forms against float32) and the CPU
baseline (the oracle's float32 step on the host cores, and the voxelizer).

Prints ONE JSON line.  `python bench.py --gpus N` with N > 1 exits: this
build runs on a single GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X published HBM3E bandwidth
BYTES_PER_CELL = {"fp32": 152, "fp32exact": 152, "fp16": 76, "fp16sr": 76}  # 19 read + 19 written
GRID = (320, 160, 128)
U_IN = 0.08
RE_REQUESTED = 1.0e5
BODY_LENGTH = 96.0
CHUNK = 100


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 20000 for the vehicle, 4000 for the sphere)")
    ap.add_argument("--warmup", type=int, default=3000,
                    help="untimed steps; from 3000 on, the first 3000 choose the launch shape")
    ap.add_argument("--ajuste", default=None,
                    help='fixed launch order as JSON, e.g. \'{"sweep": 0, "rotate": 9216}\' '
                         "(skips the launch-order autotuner)")
    ap.add_argument("--precision", choices=("fp32", "fp32exact", "fp16", "fp16sr"),
                    default="fp32")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="budget of the CPU-baseline sample (0 disables it)")
    ap.add_argument("--workload", choices=("vehicle", "sphere"), default="vehicle",
                    help="vehicle: 320x160x128 (the headline); sphere: the 160x80x64 "
                         "icosphere case, 4000 steps")
    ap.add_argument("--no-precision-check", dest="precision_check", action="store_false",
                    help="skip the float16-vs-float32 Cd/St check after the timed region")
    ap.add_argument("--campos", default=None,
                    help="sphere workload: write the final macro field (VTK) here")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 4000 if args.workload == "sphere" else 20000
    return args


def traffic_from_profiles(precision: str):
    """HBM bytes per step-kernel launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), read
    from the committed summary; None if it is absent."""
    f = ROOT / "profiles" / f"step_{precision}_320x160x128_pmc.md"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def build_case(precision: str):
    from cfd__amd import meshes, tunel
    from cfd__amd.units import Units
    from cfd__amd.voxel import voxelizar

    nx, ny, nz = GRID
    g = {"nx": nx, "ny": ny, "nz": nz}
    pos, idx, info = meshes.vehicle(seed=7)
    voxelizar(pos, idx, g, banda=1.5, selarFrestas=True, comSdf=False)   # device warm-up
    vox = voxelizar(pos, idx, g, banda=1.5, selarFrestas=True, comSdf=False)
    tipo, esponja, cont = tunel.montarTipos(g, vox["tipo"], esteira=False)
    area = tunel.areaFrontalCorpo(vox["tipo"], g)
    # Physical scale chosen so that Re_physical = 1e5 for the 96-cell body.
    length_m = 4.5
    speed = RE_REQUESTED * 1.506e-5 / length_m
    units = Units(lengthM=length_m, speedMs=speed, cellsPerLength=BODY_LENGTH, uLb=U_IN,
                  rollingRoad=False, lesCs=0.1)
    return dict(grid=g, tipo=tipo, esponja=esponja, area=area, units=units, vox=vox,
                mesh=info, cont=cont, pos=pos, idx=idx)


RAMP_STEPS = int(math.floor(GRID[0] / 0.05 + 0.5))   # the app's ramp length (nx / 0.05)
WINDOW = 5000                                          # Cd/Cl averaging window (last steps)


def run_gpu(case, precision: str, steps: int, warmup: int, ajuste=None):
    from cfd__amd.forces import JanelaCd, coeficientes
    from cfd__amd.solver import SolverMI355X

    g, units = case["grid"], case["units"]
    s = SolverMI355X.criar(g, precisao=precision)
    # fixed floor: the app feeds a 1/7-power inlet boundary layer of
    # round(0.12 nz) cells (index.html:727-734)
    cfg = dict(omegaPlus=units.omegaPlus, lesCs=0.1, esponja=case["esponja"],
               beltU=[0.0, 0.0, 0.0], espessuraCL=float(math.floor(g["nz"] * 0.12 + 0.5)))

    def aplicar(f):
        s.configurar(inletU=[U_IN * f, 0.0, 0.0], **cfg)

    s.definirTipos(case["tipo"])
    aplicar(1.0)
    s.calibrarRepouso()
    s.rampa(RAMP_STEPS, aplicar)
    s.inicializar()

    # warm-up (untimed); when it is long enough, its first steps choose the
    # launch shape in calls of CHUNK steps, as the loop below takes them
    # (results and the ramp are bit-identical in every shape)
    done = 0
    tune_steps = SolverMI355X.passosAutoajuste(bloco=CHUNK)
    if ajuste is not None:
        s.ajuste = ajuste
        origem = "fixed (--ajuste)"
    elif warmup >= tune_steps:
        s.autoajustarPasso(bloco=CHUNK)
        done = tune_steps
        origem = "autotuned in the warm-up"
    else:
        origem = "library default for this population set (cfd_create)"
    ordem = {k: s.ajuste[k] for k in ("block", "remap", "sweep", "rotate")}
    while done < warmup:
        n = min(CHUNK, warmup - done)
        s.passo(n)
        done += n
    s.sincronizar()

    cd_win, cl_win = JanelaCd(10 ** 6), JanelaCd(10 ** 6)
    sample_from = max(0, steps - WINDOW)
    first_sample = None
    kernel_ms = 0.0
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(CHUNK, steps - done)
        kernel_ms += s.passoCronometrado(n)
        done += n
        if done > sample_from:
            f = s.medirForcas()
            co = coeficientes(f, units, case["area"], g, corrigirBloqueio=True, theta=2.5)
            cd_win.adicionar(co["cd"])
            cl_win.adicionar(co["cl"])
            if first_sample is None:
                first_sample = warmup + done
    s.sincronizar()
    wall = time.perf_counter() - t0
    mass, nonfinite = s.massa()
    s.destruir()
    return dict(wall=wall, kernel_ms=kernel_ms, ordem=ordem, origem=origem,
                cd=cd_win.media, cd_std=cd_win.desvio,
                cl=cl_win.media, cl_std=cl_win.desvio, n_samples=cd_win.n,
                first_sample=first_sample, last_step=warmup + steps, mass=mass,
                nonfinite=nonfinite)


def _finite(x):
    return x if x is not None and math.isfinite(x) else None


def precision_check():
    """Cd (and St) of the float16 storage forms against plain float32 on the
    reference's validation cases (validacao.js:151-205: sphere Re 100 at
    240x120x96, cylinder Re 100 at 480x8x320 with its shedding) and on the
    160x80x64 sphere workload; every run on the GPU, after the timed region."""
    from cfd__amd.caso import rodar_caso
    from cfd__amd.validacao import CASOS

    out = {}
    for cid in ("esfera-re100", "cilindro-re100"):
        caso = next(c for c in CASOS if c["id"] == cid)
        runs = {}
        for prec in ("fp32", "fp16", "fp16sr"):
            r = rodar_caso(caso, precisao=prec)
            runs[prec] = {"cd": r["cd"], "st": r.get("st"), "nonfinite": r["naoFinitos"],
                          "steps": r["passos"]}
        out[cid] = runs
    c = sphere_case()
    runs = {}
    for prec in ("fp32", "fp16", "fp16sr"):
        cd, mass, bad, _, _, _ = run_sphere(c, prec, 4000, 0)
        runs[prec] = {"cd": cd, "mass": mass, "nonfinite": bad, "steps": 4000}
    out["esfera-d16-160x80x64"] = runs
    for runs in out.values():
        ref = runs["fp32"]
        for prec in ("fp16", "fp16sr"):
            r = runs[prec]
            r["cd_err_vs_fp32"] = _finite((r["cd"] - ref["cd"]) / ref["cd"])
            if r.get("st") is not None and ref.get("st"):
                r["st_err_vs_fp32"] = _finite((r["st"] - ref["st"]) / ref["st"])
    return out


def reference_voxelizer_ms():
    """The reference's own voxelizar() (src/geom/voxel.js) on this 200k-triangle
    vehicle, timed under Node when the golden fixtures were written."""
    f = ROOT / "tests" / "golden" / "voxel_golden.json"
    try:
        case = json.loads(f.read_text())["cases"]["vehicle_200k_320x160x128"]
        return case.get("ms_reference")
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(case, budget_s: float):
    """The oracle's float32 step (OpenMP, all host cores) on the same cell
    types, from the initialized state, timed for about budget_s seconds."""
    import oracle

    nx, ny, nz = GRID
    par = oracle.params_array(omegaPlus=case["units"].omegaPlus, lesCs=0.1,
                              inletU=[U_IN, 0.0, 0.0], esponja=case["esponja"],
                              espessuraCL=float(math.floor(nz * 0.12 + 0.5)), nx=nx)
    a = oracle.init(GRID, case["tipo"], par)
    b = a.copy()
    lib = oracle.lib()
    t = np.ascontiguousarray(case["tipo"], dtype=np.uint32)
    P = oracle._p
    lib.orc_step_f32(nx, ny, nz, P(a), P(b), P(t), P(par), None)  # warm-up
    a, b = b, a
    n = 0
    t0 = time.perf_counter()
    while True:
        lib.orc_step_f32(nx, ny, nz, P(a), P(b), P(t), P(par), None)
        a, b = b, a
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 1000:
            break
    cells = nx * ny * nz
    # the voxelizer beside it: the oracle's restatement, single-threaded like
    # the reference's JavaScript, on the same 200k-triangle mesh (it always
    # computes the signed distance, which the GPU run above skips)
    t1 = time.perf_counter()
    oracle.voxelize(case["pos"], case["idx"], GRID, 1.5, True)
    vox_ms = (time.perf_counter() - t1) * 1e3
    return {"value": cells * n / el / 1e6, "unit": "MLUPS", "cores": oracle.threads(),
            "kind": "port",
            "sample": f"oracle float32 step (oracle/lbm_oracle.c, OpenMP) on the same "
                      f"320x160x128 vehicle cell types, {n} steps after 1 warm-up step, "
                      f"{el:.1f} s",
            "voxelizer": {"ms": round(vox_ms, 1), "cores": 1, "kind": "port",
                          "sample": "oracle voxelizer (single thread, with the signed "
                                    "distance) on the same 200k-triangle vehicle, 320x160x128",
                          "reference": {
                              "ms": _r1(reference_voxelizer_ms()), "cores": 1,
                              "kind": "reference",
                              "sample": "the reference's own voxelizar() (src/geom/voxel.js, "
                                        "Node 12, one thread) on this 200k-triangle vehicle "
                                        "at 320x160x128, timed when "
                                        "tests/golden/voxel_golden.json was written, where "
                                        "the reference's sources were present (they are "
                                        "not needed, and not read, at run time)"}}}


def _r1(x):
    return round(x, 1) if x is not None else None


def sphere_case():
    """160x80x64, seeded 20 480-triangle icosphere D = 16 at (48, 40, 32),
    free-slip sides, top and floor (the body in free air), u = 0.05, nu for
    Re = 100 (omega 1.9084, just above the reference's 1.90 ceiling), Cs 0.1."""
    from cfd__amd import meshes, tunel
    from cfd__amd.units import Units
    from cfd__amd.voxel import voxelizar

    nx, ny, nz = 160, 80, 64
    g = {"nx": nx, "ny": ny, "nz": nz}
    pos, idx = meshes.icosphere(5, 16.0, (48.0, 40.0, 32.0), 1)
    voxelizar(pos, idx, g, banda=1.5, selarFrestas=True, comSdf=False)   # device warm-up
    t0 = time.perf_counter()
    vox = voxelizar(pos, idx, g, banda=1.5, selarFrestas=True, comSdf=False)
    vox_ms = (time.perf_counter() - t0) * 1e3
    tipo, esponja, cont = tunel.montarTipos(g, vox["tipo"], esteira=False)
    tipo.reshape(nz, ny, nx)[0, :, 1:nx - 1] = 6
    nu = 0.05 * 16 / 100
    units = Units(lengthM=0.016, speedMs=0.1, cellsPerLength=16, uLb=0.05, lesCs=0.1)
    return dict(grid=g, tipo=tipo, esponja=esponja, area=tunel.areaFrontalCorpo(vox["tipo"], g),
                omega=1 / (3 * nu + 0.5), u=0.05, units=units, n_tri=len(idx) // 3,
                voxel_ms=vox_ms)


def run_sphere(c, precision: str, steps: int, warmup: int, campos=None):
    """The sphere workload on the GPU: Cd over the last 1000 steps (a sample
    every 100), final mass and non-finite count, wall and kernel time."""
    from cfd__amd.forces import JanelaCd, coeficientes
    from cfd__amd.solver import SolverMI355X

    g = c["grid"]
    cfg = dict(omegaPlus=c["omega"], lesCs=0.1, inletU=[c["u"], 0, 0], esponja=c["esponja"])
    s = SolverMI355X.criar(g, precisao=precision)
    try:
        s.definirTipos(c["tipo"])
        s.configurar(**cfg)
        s.calibrarRepouso()
        s.inicializar()
        if warmup > 0:
            s.passo(warmup)
            s.sincronizar()
            s.inicializar()          # the timed run starts from the initial state again
        jan = JanelaCd(10 ** 6)
        kernel_ms = 0.0
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(CHUNK, steps - done)
            kernel_ms += s.passoCronometrado(n)
            done += n
            if done > steps - 1000:
                co = coeficientes(s.medirForcas(), c["units"], c["area"], g,
                                  corrigirBloqueio=True, theta=2.5)
                jan.adicionar(co["cd"])
        s.sincronizar()
        wall = time.perf_counter() - t0
        if campos:
            s.salvarCampo(campos, c["u"])
        mass, bad = s.massa()
    finally:
        s.destruir()
    return jan.media, mass, bad, wall, kernel_ms, jan.desvio


def main_sphere(args):
    c = sphere_case()
    g = c["grid"]
    cells = g["nx"] * g["ny"] * g["nz"]
    cfg = dict(omegaPlus=c["omega"], lesCs=0.1, inletU=[c["u"], 0, 0], esponja=c["esponja"])
    cd, mass, bad, wall, kernel_ms, cd_std = run_sphere(c, args.precision, args.steps,
                                                        max(args.warmup, 0), args.campos)
    cpu = None
    if args.cpu_seconds > 0:
        import oracle
        par = oracle.params_array(nx=g["nx"], **cfg)
        grid = (g["nx"], g["ny"], g["nz"])
        a = oracle.init(grid, c["tipo"], par)
        b = a.copy()
        lib, P = oracle.lib(), oracle._p
        t = np.ascontiguousarray(c["tipo"], dtype=np.uint32)
        n = 0
        t1 = time.perf_counter()
        while time.perf_counter() - t1 < args.cpu_seconds and n < 4000:
            lib.orc_step_f32(*grid, P(a), P(b), P(t), P(par), None)
            a, b = b, a
            n += 1
        el = time.perf_counter() - t1
        cpu = {"value": cells * n / el / 1e6, "unit": "MLUPS", "cores": oracle.threads(),
               "kind": "port", "sample": f"oracle float32 step, same case, {n} steps, {el:.1f} s"}
    kernel_avg_s = kernel_ms / 1e3 / args.steps
    bpc = BYTES_PER_CELL[args.precision]
    print(json.dumps({
        "metric": "MLUPS (D3Q19 TRT-LES, 0.82M cells, sphere)", "unit": "MLUPS",
        "value": round(cells * args.steps / wall / 1e6, 1), "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision.startswith("fp32") else "f16-storage/f32-collide",
        "data": f"synthetic (seeded {c['n_tri']}-triangle icosphere, GPU-voxelized)",
        "config": {"workload": f"plain D3Q19 TRT(3/16)+Smagorinsky(0.1) {args.precision} "
                               "sphere D16 160x80x64, u_in=0.05, Re=100",
                   "omega_plus": c["omega"], "cells": cells},
        "roofline": {"bound": "hbm", "achieved": round(bpc * cells / kernel_avg_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bpc * cells / kernel_avg_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_kernel_us": round(kernel_avg_s * 1e6, 2),
                     "note": "the 125 MB population set fits the 256 MB last-level cache"},
        "cpu_baseline": cpu,
        "results": {"cd_mean_last_1000": _finite(cd), "cd_std": _finite(cd_std),
                    "frontal_area_cells": c["area"], "mass_final": mass, "nonfinite_cells": bad,
                    "voxelize_ms": round(c["voxel_ms"], 2),
                    "campos": args.campos},
    }))


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sphere":
        return main_sphere(args)
    from cfd__amd import _lib

    precision = args.precision
    case = build_case(precision)
    nx, ny, nz = GRID
    cells = nx * ny * nz
    r = run_gpu(case, precision, args.steps, args.warmup,
                json.loads(args.ajuste) if args.ajuste else None)
    mlups = cells * args.steps / r["wall"] / 1e6
    kernel_avg_s = r["kernel_ms"] / 1e3 / args.steps
    bpc = BYTES_PER_CELL[precision]
    achieved = bpc * cells / kernel_avg_s / 1e9
    copy_gbs, _ = _lib.copy_bandwidth(1 << 30, 20)
    prec = precision_check() if args.precision_check else None
    cpu = cpu_baseline(case, args.cpu_seconds) if args.cpu_seconds > 0 else None
    u = case["units"]
    line = {
        "metric": "MLUPS (D3Q19 TRT-LES, 6.55M cells)",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if precision.startswith("fp32") else "f16-storage/f32-collide",
        "data": "synthetic (seeded 200k-triangle vehicle, GPU-voxelized)",
        "config": {
            "workload": f"plain D3Q19 TRT(3/16)+Smagorinsky(0.1) {precision} vehicle "
                        f"{nx}x{ny}x{nz}, u_in={U_IN}",
            "grid": [nx, ny, nz], "cells": cells, "precision": precision,
            "omega_plus_used": u.omegaPlus, "re_lattice_used": u.reLattice,
            "omega_plus_requested": 1 / (3 * U_IN * BODY_LENGTH / RE_REQUESTED + 0.5),
            "re_requested": RE_REQUESTED, "floor": "PAREDE (no-slip)",
            "inlet_boundary_layer_cells": math.floor(nz * 0.12 + 0.5),
            "parallelism": "single GPU",
            "launch_order": r["ordem"], "launch_order_source": r["origem"],
        },
        "hbm_frac_of_peak": round(achieved / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(precision),
            "kernel": "k_step (fused pull-stream + TRT + LES + boundaries)",
            "bytes_per_cell": bpc,
            "avg_kernel_us": round(kernel_avg_s * 1e6, 2),
        },
        "cpu_baseline": cpu,
        "results": {
            "cd_mean": _finite(r["cd"]), "cd_std": _finite(r["cd_std"]),
            "cl_mean": _finite(r["cl"]), "cl_std": _finite(r["cl_std"]),
            "force_window": {
                "samples": r["n_samples"], "first_step": r["first_sample"],
                "last_step": r["last_step"], "every": CHUNK,
                "ramp_steps": RAMP_STEPS,
                # the configured measurement: mean of the last 5000 of 20000 steps
                "as_configured": r["last_step"] >= 20000 and r["first_sample"] is not None
                and r["last_step"] - r["first_sample"] + CHUNK >= WINDOW,
                "transient": r["first_sample"] is None or r["first_sample"] <= RAMP_STEPS,
            },
            "frontal_area_cells": case["area"], "solid_cells": case["cont"]["solido"],
            "voxelize_ms": case["vox"]["stats"]["msTotal"],
            "mass_final": r["mass"], "nonfinite_cells": r["nonfinite"],
            "copy_kernel_gbs": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
        },
        "precision": prec,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
