"""Physical <-> lattice units (src/core/units.js), float64."""
from __future__ import annotations

import math

from .lattice import CS2, nuFromOmega, omegaFromNu

AR = {"rho": 1.204, "mu": 1.813e-5, "nu": 1.506e-5, "c": 343.2}  # units.js:53-58
OMEGA_MAX = 1.90            # units.js:91
OMEGA_MAX_PISO_PARADO = 1.90
NU_MIN = nuFromOmega(OMEGA_MAX)
U_LB_PADRAO = 0.05
U_LB_MAX = 0.1


class Units:
    """units.js:110-171: the viscosity is clamped at nu(OMEGA_MAX)."""

    def __init__(self, lengthM: float, speedMs: float, cellsPerLength: float,
                 uLb: float = U_LB_PADRAO, rollingRoad: bool = True, fluid=None,
                 lesCs: float = 0.1):
        if not (lengthM > 0):
            raise ValueError("lengthM must be positive")
        if not (speedMs > 0):
            raise ValueError("speedMs must be positive")
        if not (cellsPerLength >= 8):
            raise ValueError(f"cellsPerLength = {cellsPerLength}: below ~8 cells the body "
                             "has no shape. Increase the resolution.")
        if not (0 < uLb <= U_LB_MAX):
            raise ValueError(f"uLb = {uLb} outside (0, {U_LB_MAX}]")
        self.fluid = dict(fluid or AR)
        self.lengthM = lengthM
        self.speedMs = speedMs
        self.cellsPerLength = int(math.floor(cellsPerLength + 0.5))
        self.uLb = uLb
        self.rollingRoad = bool(rollingRoad)
        self.lesCs = lesCs
        self.dx = lengthM / self.cellsPerLength
        self.dt = uLb * self.dx / speedMs
        self.rePhysical = speedMs * lengthM / self.fluid["nu"]
        self.nuLbExact = uLb * self.cellsPerLength / self.rePhysical
        self.omegaMax = OMEGA_MAX if self.rollingRoad else OMEGA_MAX_PISO_PARADO
        self.nuLbMin = nuFromOmega(self.omegaMax)
        self.nuLb = max(self.nuLbExact, self.nuLbMin)
        self.omegaPlus = omegaFromNu(self.nuLb)
        self.reLattice = uLb * self.cellsPerLength / self.nuLb
        self.machLattice = uLb / math.sqrt(CS2)
        self.machPhysical = speedMs / self.fluid["c"]

    @property
    def resolved(self) -> bool:
        return self.nuLbExact >= self.nuLbMin

    @property
    def reRatio(self) -> float:
        return self.rePhysical / self.reLattice

    def velToLb(self, v):
        return v * self.uLb / self.speedMs

    def velToSi(self, v):
        return v * self.speedMs / self.uLb

    def lenToLb(self, x):
        return x / self.dx

    def lenToSi(self, x):
        return x * self.dx

    def timeToSi(self, steps):
        return steps * self.dt

    def pressureToSi(self, delta):
        return CS2 * delta * self.fluid["rho"] * (self.dx / self.dt) ** 2

    def forceToSi(self, f):
        return f * self.fluid["rho"] * self.dx ** 4 / self.dt ** 2

    @property
    def dynamicPressure(self) -> float:
        return 0.5 * self.fluid["rho"] * self.speedMs ** 2


def stepsForSeconds(units: Units, seconds: float) -> int:
    return math.ceil(seconds / units.dt)


def stepsForFlowthrough(units: Units, n: float = 1) -> int:
    """units.js:259-261"""
    return math.ceil(n * units.cellsPerLength / units.uLb)
