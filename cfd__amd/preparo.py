"""From the mesh as exported to the mesh placed in the tunnel
(src/geom/prepare.js): which axis is the flow, which is up, what the unit is,
then permutation, yaw, scale and placement in lattice units.

Positions stay float32 between stages with float64 arithmetic inside each
stage, as the reference's Float32Array code does, so a mesh placed here
voxelizes to the same cells.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh_io import extentos


def _js_round(x: float) -> int:
    return int(math.floor(x + 0.5))


def ocupacao(pos: np.ndarray, ext: dict, K: int = 24) -> tuple[np.ndarray, int]:
    """Coarse K^3 occupancy over the bounding box (prepare.js:44-57)."""
    p = np.asarray(pos, dtype=np.float32).reshape(-1, 3).astype(np.float64)
    ok = np.isfinite(p.sum(axis=1))
    p = p[ok]
    tam = np.array(ext["tamanho"], dtype=np.float64)
    inv = np.where(tam > 0, K / np.where(tam > 0, tam, 1.0), 0.0)
    lo = np.array(ext["min"], dtype=np.float64)
    idx = np.trunc((p - lo) * inv)
    idx = np.clip(idx, 0, K - 1).astype(np.int64)
    oc = np.zeros(K * K * K, dtype=np.uint8)
    oc[(idx[:, 2] * K + idx[:, 1]) * K + idx[:, 0]] = 1
    return oc, K


def simetria(oc: np.ndarray, K: int, eixo: int) -> float:
    """Fraction of occupied voxels whose mirror across axis `eixo` is
    occupied too (prepare.js:59-75)."""
    g = oc.reshape(K, K, K)                       # [c][b][a] = z, y, x
    m = np.flip(g, axis=2 - eixo)
    ocupados = int(g.sum())
    casados = int((g & m).sum())
    return casados / ocupados if ocupados else 0.0


def palpitarEixos(pos, ext: dict) -> dict:
    """Flow axis = the longest extent; of the other two the more symmetric is
    the lateral one, the other vertical (prepare.js:82-111)."""
    tam = ext["tamanho"]
    ordem_tam = sorted([0, 1, 2], key=lambda a: -tam[a])   # stable, as Array.sort
    fluxo = ordem_tam[0]
    r0, r1 = [a for a in (0, 1, 2) if a != fluxo]
    oc, K = ocupacao(pos, ext)
    sim = [simetria(oc, K, a) for a in range(3)]
    lateral = r0 if sim[r0] >= sim[r1] else r1
    vertical = r1 if lateral == r0 else r0
    return {"ordem": [fluxo, lateral, vertical], "simetrias": sim,
            "confianca": abs(sim[r0] - sim[r1]), "talvezInvertido": None}


def palpitarSentido(pos, ordem, ext: dict) -> dict:
    """Nose low, tail high: compare the top of the silhouette over the first
    and last 20 % of the length (prepare.js:138-172)."""
    ex, ez = ordem[0], ordem[2]
    x0 = ext["min"][ex]
    comprimento = max(ext["tamanho"][ex], 1e-12)
    zbase = ext["min"][ez]
    N = 12
    p = np.asarray(pos, dtype=np.float32).reshape(-1, 3).astype(np.float64)
    x, z = p[:, ex], p[:, ez]
    ok = np.isfinite(x + z)
    k = np.clip(np.floor((x[ok] - x0) / comprimento * N), 0, N - 1).astype(np.int64)
    h = z[ok] - zbase
    alturas = np.zeros(N)
    np.maximum.at(alturas, k, h)
    alturas = np.maximum(alturas, 0.0)
    n = _js_round(N * 0.2)
    frente = float(alturas[:n].sum() / n)
    tras = float(alturas[N - n:].sum() / n)
    escala = max(frente, tras, 1e-12)
    return {"inverter": frente > tras, "alturaFrente": frente, "alturaTras": tras,
            "confianca": abs(tras - frente) / escala}


def palpitarUnidade(comprimento: float) -> dict:
    """Scale to metres from a plausible vehicle length (prepare.js:180-212)."""
    candidatos = [
        {"fator": 1, "nome": "metros", "tipico": [0.5, 30]},
        {"fator": 0.01, "nome": "centímetros", "tipico": [50, 3000]},
        {"fator": 0.001, "nome": "milímetros", "tipico": [500, 30000]},
        {"fator": 0.0254, "nome": "polegadas", "tipico": [20, 1200]},
    ]
    for c in candidatos:
        if c["tipico"][0] <= comprimento <= c["tipico"][1]:
            return {**c, "normalizado": False, "confianca": "plausível"}
    if 0.2 < comprimento < 2.5:
        return {"fator": 1, "nome": "normalizado", "normalizado": True, "confianca": "nenhuma",
                "aviso": f"largest dimension = {comprimento:.3f}; the model looks normalized. "
                         "Give the real body length — Re and Cd depend on it."}
    return {"fator": 1, "nome": "desconhecida", "normalizado": True, "confianca": "nenhuma",
            "aviso": f"largest dimension = {comprimento:.4g}, outside any typical range. "
                     "Give the real body length."}


def colocar(positions, ordem, grade: dict, celulasPorComprimento: float,
            espelhar=(False, False, False), fracaoEntrada: float = 0.28,
            assentarNoPiso: bool = True, folgaPiso: float = 1, guinadaGraus: float = 0) -> dict:
    """Permute (+ mirror), yaw about the vertical, scale the length to
    `celulasPorComprimento` cells, put the nose at fracaoEntrada * nx, centre
    in y, rest on the floor (prepare.js:226-302).  Returns a new array."""
    src = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
    saida = np.empty_like(src)
    for a in range(3):
        v = src[:, ordem[a]]
        saida[:, a] = -v if espelhar[a] else v
    if guinadaGraus:
        t = guinadaGraus * math.pi / 180
        cs, sn = math.cos(t), math.sin(t)
        e0 = extentos(saida)
        cx, cy = e0["centro"][0], e0["centro"][1]
        x = saida[:, 0].astype(np.float64) - cx
        y = saida[:, 1].astype(np.float64) - cy
        saida[:, 0] = (cx + x * cs - y * sn).astype(np.float32)
        saida[:, 1] = (cy + x * sn + y * cs).astype(np.float32)
    ext = extentos(saida)
    comprimento = max(ext["tamanho"][0], 1e-12)
    s = celulasPorComprimento / comprimento
    alvo = [grade["nx"] * fracaoEntrada - ext["min"][0] * s,
            grade["ny"] * 0.5 - ext["centro"][1] * s,
            (folgaPiso - ext["min"][2] * s) if assentarNoPiso
            else grade["nz"] * 0.5 - ext["centro"][2] * s]
    for a in range(3):
        saida[:, a] = (saida[:, a].astype(np.float64) * s + alvo[a]).astype(np.float32)
    fim = extentos(saida)
    bloqueio = (fim["tamanho"][1] * fim["tamanho"][2]) / (grade["ny"] * grade["nz"])
    avisos = []
    for a, nome in ((0, "x"), (1, "y"), (2, "z")):
        lim = (grade["nx"], grade["ny"], grade["nz"])[a]
        if fim["min"][a] < 0 or fim["max"][a] > lim:
            avisos.append(f"the body leaves the domain in {nome} "
                          f"({fim['min'][a]:.1f}…{fim['max'][a]:.1f} of 0…{lim})")
    if bloqueio > 0.10:
        avisos.append(f"blockage {bloqueio * 100:.1f}% — above 10% the Maskell correction "
                      "stops being reliable; use a larger domain")
    return {"positions": saida.reshape(-1), "escala": s, "extentos": fim,
            "bloqueioCaixa": bloqueio, "avisos": avisos}


def preparar(malha: dict, grade: dict, opcoes: dict | None = None) -> dict:
    """Raw positions -> what the voxelizer consumes, every guess explicit and
    overridable (prepare.js:304-353)."""
    o = dict(opcoes or {})
    ext = extentos(malha["positions"])
    if o.get("ordem") is not None:
        eixos = {"ordem": list(o["ordem"]), "confianca": 1, "simetrias": None, "manual": True}
    else:
        eixos = palpitarEixos(malha["positions"], ext)
    comp_arq = ext["tamanho"][eixos["ordem"][0]]
    unidade = o.get("unidade") or palpitarUnidade(comp_arq)
    comprimento_m = o.get("comprimentoM", comp_arq * unidade["fator"])
    sentido = palpitarSentido(malha["positions"], eixos["ordem"], ext)
    inverter = o["inverterSentido"] if o.get("inverterSentido") is not None else sentido["inverter"]
    # a half turn about the vertical (x -> -x, y -> -y), not a mirror in x
    espelhar = o.get("espelhar") or ([True, True, False] if inverter else [False, False, False])
    kw = {k: o[k] for k in ("fracaoEntrada", "assentarNoPiso", "folgaPiso", "guinadaGraus")
          if o.get(k) is not None}
    col = colocar(malha["positions"], eixos["ordem"], grade,
                  o.get("celulasPorComprimento") or _js_round(grade["nx"] * 0.34),
                  espelhar=espelhar, **kw)
    avisos = list(malha.get("avisos") or []) + col["avisos"]
    if unidade.get("aviso"):
        avisos.append(unidade["aviso"])
    if sentido["confianca"] < 0.12:
        avisos.append(f"front and rear almost equally high ({sentido['alturaFrente']:.2f} vs "
                      f"{sentido['alturaTras']:.2f}): the direction is a guess. Check that the "
                      "wind meets the nose and flip it if not.")
    return {**col, "indices": malha["indices"], "eixos": eixos,
            "sentido": {**sentido, "aplicado": inverter,
                        "manual": o.get("inverterSentido") is not None},
            "unidade": unidade, "comprimentoM": comprimento_m, "extentosOriginais": ext,
            "avisos": avisos}
