"""Tunnel assembly (src/core/tunel.js) with numpy.

Cell types follow CELULA of src/core/emit/ir.js:71-80.
"""
from __future__ import annotations

import numpy as np

CELULA = {
    "FLUIDO": 0, "SOLIDO": 1, "ENTRADA": 2, "SAIDA": 3,
    "SOLIDO_MOVEL": 4, "ESPELHO_Y": 5, "ESPELHO_Z": 6, "PAREDE": 7,
}
VOXEL = {"FLUIDO": 0, "CASCA": 1, "DENTRO": 2}  # src/geom/voxel.js:155-159


def _grid(grade):
    return int(grade["nx"]), int(grade["ny"]), int(grade["nz"])


def _js_round(x: float) -> int:
    return int(np.floor(x + 0.5))


def montarTipos(grade, corpo=None, esteira: bool = True, fracaoEsponja: float = 0.18):
    """Body first, then the tunnel walls over it (tunel.js:39-108).

    Returns (tipo uint32[N], esponja {inicio, comprimento}, contagem)."""
    nx, ny, nz = _grid(grade)
    N = nx * ny * nz
    tipo = np.zeros((nz, ny, nx), dtype=np.uint32)
    if corpo is not None:
        c = np.asarray(corpo).reshape(-1)
        if c.size != N:
            raise ValueError(f"body field with {c.size}, expected {N}")
        tipo.reshape(-1)[c != VOXEL["FLUIDO"]] = CELULA["SOLIDO"]
    piso = CELULA["SOLIDO_MOVEL"] if esteira else CELULA["PAREDE"]
    tipo[:, :, 0] = CELULA["ENTRADA"]
    tipo[:, :, nx - 1] = CELULA["SAIDA"]
    tipo[:, 0, :] = CELULA["ESPELHO_Y"]
    tipo[:, ny - 1, :] = CELULA["ESPELHO_Y"]
    tipo[0, :, :] = piso
    tipo[nz - 1, :, :] = CELULA["ESPELHO_Z"]
    # inlet / outlet win the corners (tunel.js:85-90)
    tipo[:, :, 0] = CELULA["ENTRADA"]
    tipo[:, :, nx - 1] = CELULA["SAIDA"]

    comprimento = max(4, _js_round(nx * fracaoEsponja))
    esponja = {"inicio": nx - 1 - comprimento, "comprimento": comprimento}
    flat = tipo.reshape(-1)
    cnt = np.bincount(flat, minlength=8)
    contagem = {
        "fluido": int(cnt[0]), "solido": int(cnt[1]), "solidoMovel": int(cnt[4]),
        "entrada": int(cnt[2]), "saida": int(cnt[3]),
        "espelho": int(N - cnt[0] - cnt[1] - cnt[4] - cnt[2] - cnt[3]),
    }
    return flat.copy(), esponja, contagem


def areaFrontalCorpo(corpo, grade) -> int:
    """Body shadow on the y-z plane, excluding the boundary rows (tunel.js:118-133)."""
    nx, ny, nz = _grid(grade)
    c = np.asarray(corpo).reshape(nz, ny, nx)
    inner = c[1:nz - 1, 1:ny - 1, 1:nx - 1] != VOXEL["FLUIDO"]
    return int(inner.any(axis=2).sum())


def areaFrontal(tipo, grade) -> int:
    """Shadow of any non-fluid voxel (src/geom/voxel.js:443-457)."""
    nx, ny, nz = _grid(grade)
    c = np.asarray(tipo).reshape(nz, ny, nx)
    return int((c != VOXEL["FLUIDO"]).any(axis=2).sum())


def corpoAnalitico(grade, forma) -> np.ndarray:
    """Sphere or y-spanning cylinder tested at cell centres (tunel.js:144-168)."""
    nx, ny, nz = _grid(grade)
    f = forma["tipo"]
    cx, cy, cz, raio = (float(forma[k]) for k in ("cx", "cy", "cz", "raio"))
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nx, dtype=np.float64), indexing="ij")
    dx, dy, dz = x + 0.5 - cx, y + 0.5 - cy, z + 0.5 - cz
    if f == "esfera":
        dentro = dx * dx + dy * dy + dz * dz <= raio * raio
    elif f == "cilindro":
        dentro = dx * dx + dz * dz <= raio * raio
    else:
        raise ValueError(f"unknown shape: {f}")
    corpo = np.zeros(nx * ny * nz, dtype=np.uint8)
    corpo[dentro.reshape(-1)] = VOXEL["DENTRO"]
    return corpo
