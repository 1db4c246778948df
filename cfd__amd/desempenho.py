"""The reference's performance page (tests/desempenho.html:58-82) on this
solver: MLUPS per grid preset of the app (src/backend/caps.js:70-88).

Per preset: an all-fluid periodic domain (cell types zero), omega+ 1.8,
Smagorinsky Cs 0.1, u_in (0.05, 0, 0), initialized; 40 warm-up steps and a
macro readback, then 200 steps closed by a macro readback (lerMacros forces
the synchronization, as on the page).  MLUPS = cells * steps / wall time;
"banda efetiva" = MLUPS * 152 B; steps/s and whole steps per 60 Hz frame.
The device time of the same 200 launches (HIP events) is reported beside
the page's wall-clock figure.

usage: python -m cfd__amd.desempenho [--presets minima,baixa,media,alta]
           [--passos 200] [--aquecimento 40] [--precisao fp32]
Prints one JSON object per preset.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

BYTES_POR_CELULA = 152

# caps.js:70-76 (the page measures the first four)
PRESETS = [
    {"id": "minima", "nome": "Mínima", "nx": 160, "ny": 80, "nz": 64},
    {"id": "baixa", "nome": "Baixa", "nx": 240, "ny": 120, "nz": 96},
    {"id": "media", "nome": "Média", "nx": 320, "ny": 160, "nz": 128},
    {"id": "alta", "nome": "Alta", "nx": 480, "ny": 240, "nz": 192},
    {"id": "extrema", "nome": "Extrema", "nx": 640, "ny": 320, "nz": 256},
]


def medir(preset: dict, passos: int = 200, aquecimento: int = 40,
          precisao: str = "fp32") -> dict:
    """One row of the page: wall-clock MLUPS with the readback that closes
    the timed steps, plus the device time of the same launches."""
    from .solver import SolverMI355X

    g = {k: preset[k] for k in ("nx", "ny", "nz")}
    s = SolverMI355X.criar(g, precisao=precisao)
    try:
        s.definirTipos(np.zeros(s.N, dtype=np.uint32))
        s.configurar(omegaPlus=1.8, lesCs=0.1, inletU=[0.05, 0.0, 0.0])
        s.inicializar()
        s.passo(aquecimento)
        s.lerMacros()
        t0 = time.perf_counter()
        s.passo(passos)
        s.lerMacros()
        dt = time.perf_counter() - t0
        ms = s.passoCronometrado(passos)
        m = s.lerMacros().reshape(-1, 4)
        finite = bool(np.isfinite(m).all())
    finally:
        s.destruir()
    N = g["nx"] * g["ny"] * g["nz"]
    mlups = N * passos / dt / 1e6
    mlups_dev = N * passos / (ms * 1e-3) / 1e6
    return {"preset": preset["id"], "cells": N, "precisao": precisao, "passos": passos,
            "mlups": mlups, "gbs": mlups * 1e6 * BYTES_POR_CELULA / 1e9,
            "passosPorSeg": passos / dt, "subpassosPorQuadro60Hz": int(passos / dt // 60),
            "mlups_dispositivo": mlups_dev,
            "gbs_dispositivo": mlups_dev * 1e6 * BYTES_POR_CELULA / 1e9,
            "finito": finite}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--presets", default="minima,baixa,media,alta")
    ap.add_argument("--passos", type=int, default=200)
    ap.add_argument("--aquecimento", type=int, default=40)
    ap.add_argument("--precisao", choices=("fp32", "fp16", "fp16sr"), default="fp32")
    a = ap.parse_args(argv)
    por_id = {p["id"]: p for p in PRESETS}
    for pid in a.presets.split(","):
        print(json.dumps(medir(por_id[pid], a.passos, a.aquecimento, a.precisao)), flush=True)


if __name__ == "__main__":
    main()
