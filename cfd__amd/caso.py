"""Headless validation cases: the loop of the reference's tests/validacao.html
(lines 169-272) on SolverMI355X.

rodar_caso(caso) builds the analytic body, the tunnel with a mirror floor
(free body), fixes omega = 1.90 and derives u_lb (validacao.js:126-129),
optionally yaws the inlet for a few hundred steps to seed vortex shedding,
runs the transient, then samples the momentum-exchange force every 100
steps into JanelaCd and the Cl series into strouhal().
"""
from __future__ import annotations

import math
import time

from .forces import JanelaCd, coeficientes, strouhal
from .solver import SolverMI355X
from .tunel import areaFrontalCorpo, corpoAnalitico, montarTipos
from .units import Units
from .validacao import classificar, uLbParaOmega

AMOSTRA = 100


def rodar_caso(caso: dict, precisao: str = "fp32", janela: int = 4000) -> dict:
    t0 = time.perf_counter()
    grade = caso["grade"]
    D = caso["diametroCelulas"]
    uLb = uLbParaOmega(caso["re"], D, 1.90)
    nuLb = uLb * D / caso["re"]
    omega = 1 / (nuLb / (1 / 3) + 0.5)
    units = Units(lengthM=0.01, speedMs=0.15, cellsPerLength=D, uLb=uLb, rollingRoad=False,
                  lesCs=0)
    units.nuLb, units.omegaPlus = nuLb, omega
    units.reLattice = uLb * D / nuLb

    corpo = corpoAnalitico(grade, {"tipo": caso["forma"], "cx": grade["nx"] * 0.30,
                                   "cy": grade["ny"] * 0.5, "cz": grade["nz"] * 0.5,
                                   "raio": D / 2})
    tipo, esponja, _ = montarTipos(grade, corpo, esteira=False)
    nx, ny = grade["nx"], grade["ny"]
    t3 = tipo.reshape(grade["nz"], ny, nx)
    t3[0, :, 1:nx - 1] = 6  # mirror floor: the body in free air (validacao.html:198-202)

    s = SolverMI355X.criar(grade, precisao=precisao)
    try:
        s.definirTipos(tipo)
        s.configurar(omegaPlus=omega, lesCs=0, inletU=[uLb, 0, 0], esponja=esponja)
        s.inicializar()
        af = areaFrontalCorpo(corpo, grade)
        travessia = math.ceil(nx / uLb)
        transiente = int(math.floor(caso["travessiasTransiente"] * travessia + 0.5))
        medida = int(math.floor(caso["travessiasMedida"] * travessia + 0.5))
        if caso.get("perturbacao"):
            t = caso["perturbacao"]["grausGuinada"] * math.pi / 180
            s.configurar(omegaPlus=omega, lesCs=0, esponja=esponja,
                         inletU=[uLb * math.cos(t), 0, uLb * math.sin(t)])
            s.passo(caso["perturbacao"]["passos"])
            s.configurar(omegaPlus=omega, lesCs=0, inletU=[uLb, 0, 0], esponja=esponja)
        k = 0
        while k < transiente:
            s.passo(min(2000, transiente - k))
            k += 2000
        jan = JanelaCd(janela)
        serie_cl = []
        divergiu = None
        k = 0
        while k < medida:
            s.passo(AMOSTRA)
            f = s.medirForcas()
            co = coeficientes(f, units, af, grade, corrigirBloqueio=True,
                              theta=1.0 if caso["forma"] == "cilindro" else 2.5)
            if not math.isfinite(co["cd"]):
                divergiu = s.passos
                break
            jan.adicionar(co["cd"])
            serie_cl.append(co["cl"])
            k += AMOSTRA
        massa, nao_finitos = s.massa()
    finally:
        s.destruir()
    ref = caso["referencia"]()
    out = {"id": caso["id"], "cd": jan.media, "cdDesvio": jan.desvio, "referencia": ref,
           "classificacao": classificar(jan.media, ref), "omega": omega, "uLb": uLb,
           "areaFrontal": af, "bloqueio": af / (grade["ny"] * grade["nz"]),
           "passos": transiente + medida, "divergiuEm": divergiu, "naoFinitos": nao_finitos,
           "segundos": time.perf_counter() - t0}
    if caso.get("referenciaSt"):
        st = strouhal(serie_cl, AMOSTRA, D, uLb)
        out["st"] = st["st"]
        out["stReferencia"] = caso["referenciaSt"]()
        out["stClassificacao"] = classificar(st["st"], out["stReferencia"])
    return out
