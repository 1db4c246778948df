"""Force -> coefficients and statistics (src/core/forces.js), float64."""
from __future__ import annotations

import math


def maskell(cdMedido: float, bloqueio: float, theta: float = 2.5) -> float:
    """Maskell blockage correction, identity for Cd <= 0 (forces.js:39-43)."""
    if not (bloqueio > 0) or not math.isfinite(cdMedido):
        return cdMedido
    if cdMedido <= 0:
        return cdMedido
    return cdMedido / (1 + theta * cdMedido * bloqueio)


def coeficientes(forcaLb, units, areaFrontalCelulas, grade, corrigirBloqueio: bool = True,
                 theta: float = 2.5) -> dict:
    """Cd, Cl, Cs in lattice units (forces.js:55-95)."""
    fx, fy, fz = forcaLb
    q = 0.5 * units.uLb * units.uLb
    denom = max(q * areaFrontalCelulas, 1e-30)
    cdBruto = fx / denom
    bloqueio = areaFrontalCelulas / (grade["ny"] * grade["nz"])
    cd = maskell(cdBruto, bloqueio, theta) if corrigirBloqueio else cdBruto
    out = {
        "cd": cd, "cdBruto": cdBruto, "cl": fz / denom, "cs": fy / denom,
        "bloqueio": bloqueio, "theta": theta,
        "corrigido": bool(corrigirBloqueio and cdBruto > 0),
    }
    if hasattr(units, "forceToSi"):
        out["forcaN"] = [units.forceToSi(fx), units.forceToSi(fy), units.forceToSi(fz)]
        out["areaFrontalM2"] = areaFrontalCelulas * units.dx * units.dx
        out["arrastoN"] = cd * units.dynamicPressure * areaFrontalCelulas * units.dx * units.dx
    return out


class JanelaCd:
    """Sliding window mean / sample std (forces.js:106-137)."""

    def __init__(self, capacidade: int = 600):
        self.capacidade = capacidade
        self.buf: list[float] = []
        self.i = 0

    def adicionar(self, v: float) -> None:
        if not math.isfinite(v):
            return
        if len(self.buf) < self.capacidade:
            self.buf.append(v)
        else:
            self.buf[self.i] = v
            self.i = (self.i + 1) % self.capacidade

    def limpar(self) -> None:
        self.buf.clear()
        self.i = 0

    @property
    def n(self) -> int:
        return len(self.buf)

    @property
    def media(self) -> float:
        if not self.buf:
            return math.nan
        s = 0.0
        for v in self.buf:
            s += v
        return s / len(self.buf)

    @property
    def desvio(self) -> float:
        n = len(self.buf)
        if n < 2:
            return math.nan
        m = self.media
        s = 0.0
        for v in self.buf:
            s += (v - m) ** 2
        return math.sqrt(s / (n - 1))

    @property
    def madura(self) -> bool:
        return len(self.buf) >= self.capacidade


def strouhal(serie, passosPorAmostra: int, dCelulas: float, uLb: float) -> dict:
    """St from zero crossings above 0.25 RMS (forces.js:151-177)."""
    n = len(serie)
    if n < 32:
        return {"st": math.nan, "motivo": "series too short"}
    m = 0.0
    for v in serie:
        m += v
    m /= n
    osc = [v - m for v in serie]
    acc = 0.0
    for b in osc:
        acc += b * b
    amp = math.sqrt(acc / n)
    if amp < 1e-6:
        return {"st": 0.0, "motivo": "no oscillation: steady wake"}
    limiar = 0.25 * amp
    cruz, ultimo = 0, 0
    for v in osc:
        s = 1 if v > limiar else (-1 if v < -limiar else 0)
        if s != 0 and ultimo != 0 and s != ultimo:
            cruz += 1
        if s != 0:
            ultimo = s
    if cruz < 4:
        return {"st": math.nan, "motivo": f"only {cruz} crossings; simulate longer"}
    periodos = cruz / 2
    freqLb = periodos / (n * passosPorAmostra)
    return {"st": freqLb * dCelulas / uLb, "periodos": periodos, "freqLb": freqLb,
            "amplitude": amp}
