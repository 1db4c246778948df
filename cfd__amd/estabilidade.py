"""The reference's stability page (tests/estabilidade.html:44-93) on this
solver: where the ghost modes start to blow up with a voxelized vehicle in
the 320x160x128 tunnel, the measurement behind OMEGA_MAX = 1.90
(src/core/units.js:60-91).

The page's vehicle is a Tesla GLB placed 32 cells long (preparar with
celulasPorComprimento 32); the model file is not in the reference tree, so
the body here is the seeded synthetic vehicle (meshes.vehicle) through the
same preparar -> voxelizar -> montarTipos path, rolling road on.  Per case
(omega+ in {1.98, 1.96, 1.94, 1.92, 1.88, 1.82} x Cs in {0.1, 0.16, 0}):
initialize, then 13 000 steps in chunks of 1000; after each chunk the macro
field is refreshed and read, |u| is sampled on every 11th cell, and the case
has diverged when a sample is non-finite or |u| > 0.30 (six times the free
stream).  The page runs the cases one after another on one solver; so does
this.

usage: python -m cfd__amd.estabilidade [--passos 13000] [--omegas ...] [--les ...]
Prints one JSON object per case.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

GRADE = {"nx": 320, "ny": 160, "nz": 128}
PASSOS = 13000
LIMITE_U = 0.30
OMEGAS = (1.98, 1.96, 1.94, 1.92, 1.88, 1.82)
LES = (0.1, 0.16, 0.0)


def montar(seed: int = 7, celulas: float = 32, esteira: bool = True):
    """Type field and sponge of the page's tunnel with the synthetic vehicle."""
    from . import meshes, tunel
    from .preparo import preparar
    from .voxel import voxelizar

    pos, idx, _ = meshes.vehicle(seed=seed)
    prep = preparar({"positions": pos, "indices": idx}, GRADE, {"celulasPorComprimento": celulas})
    vox = voxelizar(prep["positions"], prep["indices"], GRADE, comSdf=False)
    tipo, esponja, _ = tunel.montarTipos(GRADE, vox["tipo"], esteira=esteira)
    return tipo, esponja


def divergiu(s, limite: float = LIMITE_U, passo_amostra: int = 11) -> tuple[bool, float]:
    """The page's divergence check (estabilidade.html:82-88): refresh the
    macro field, sample |u| on every `passo_amostra`-th cell; diverged when a
    sample is non-finite or above `limite`.  Returns (diverged, max |u|),
    max |u| NaN when a sample is not finite."""
    s.atualizarMacros()
    m = s.lerMacros().reshape(-1, 4)[::passo_amostra, :3].astype(np.float64)
    u = np.sqrt((m * m).sum(axis=1))
    if not np.isfinite(u).all():
        return True, float("nan")
    max_u = float(u.max()) if u.size else 0.0
    return max_u > limite, max_u


def rodar_caso(s, omega: float, les: float, esponja: dict, passos: int = PASSOS,
               checar: int = 1000) -> dict:
    """One row of the page on an existing solver (types already set)."""
    s.configurar(omegaPlus=omega, lesCs=les, inletU=[0.05, 0.0, 0.0], beltU=[0.05, 0.0, 0.0],
                 esponja=esponja)
    s.inicializar()
    morreu, max_u = None, 0.0
    k = 0
    while k < passos and morreu is None:
        s.passo(checar)
        k += checar
        ruim, max_u = divergiu(s)
        if ruim:
            morreu = s.passos
    nu = (1.0 / 3.0) * (1.0 / omega - 0.5)
    return {"omega": omega, "les": les, "nu": nu, "morreu": morreu, "maxU": max_u,
            "passos": passos}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--passos", type=int, default=PASSOS)
    ap.add_argument("--omegas", default=",".join(str(o) for o in OMEGAS))
    ap.add_argument("--les", default=",".join(str(c) for c in LES))
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args(argv)
    from .solver import SolverMI355X

    tipo, esponja = montar(a.seed)
    s = SolverMI355X.criar(GRADE)
    try:
        s.definirTipos(tipo)
        for les in (float(c) for c in a.les.split(",")):
            for omega in (float(o) for o in a.omegas.split(",")):
                print(json.dumps(rodar_caso(s, omega, les, esponja, a.passos)), flush=True)
    finally:
        s.destruir()


if __name__ == "__main__":
    main()
