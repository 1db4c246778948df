"""D3Q19 velocity set and float64 helpers (src/core/lattice.js).

Shifted populations: the solver stores g_i = f_i - w_i, so rest is g = 0,
delta = sum g_i, rho u = sum g_i c_i (lattice.js:1-46).
"""
from __future__ import annotations

import numpy as np

Q = 19
CS2 = 1.0 / 3.0
INV_CS2 = 3.0
INV_CS4 = 9.0

# lattice.js:57-79 - opposite pairs adjacent
C = np.array([
    [0, 0, 0],
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
    [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
    [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
], dtype=np.int64)
# lattice.js:82-86
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
# lattice.js:89: opp(i) = ((i-1) ^ 1) + 1
OPP = np.array([0] + [((k ^ 1) + 1) for k in range(Q - 1)], dtype=np.int64)

# lattice.js:111-116
MAGIC = {"WALL": 3 / 16, "STABLE": 1 / 4, "THIRD_ORDER": 1 / 12, "FOURTH_ORDER": 1 / 6}


def omegaFromNu(nu: float) -> float:
    """lattice.js:121-123"""
    return 1 / (nu / CS2 + 0.5)


def nuFromOmega(omega: float) -> float:
    """lattice.js:126-128"""
    return CS2 * (1 / omega - 0.5)


def omegaMinusFromLambda(omegaPlus: float, lam: float) -> float:
    """lattice.js:131-134"""
    tau_minus_half = lam / (1 / omegaPlus - 0.5)
    return 1 / (tau_minus_half + 0.5)


def lambdaFromOmegas(omegaPlus: float, omegaMinus: float) -> float:
    """lattice.js:137-139"""
    return (1 / omegaPlus - 0.5) * (1 / omegaMinus - 0.5)


def espelharIdx(i: int, eixo: int) -> int:
    """Direction with component `eixo` negated (ir.js:86-93)."""
    c = C[i].copy()
    c[eixo] = -c[eixo]
    for j in range(Q):
        if (C[j] == c).all():
            return j
    raise ValueError(f"espelharIdx: no pair for {i} on axis {eixo}")


def equilibrium(delta, u, out=None):
    """Shifted equilibrium in float64 (lattice.js:150-160).

    Scalars give a (19,) array; arrays of shape S give (19, *S)."""
    delta = np.asarray(delta, dtype=np.float64)
    ux, uy, uz = (np.asarray(v, dtype=np.float64) for v in u)
    rho = 1 + delta
    uu = ux * ux + uy * uy + uz * uz
    g = out if out is not None else np.empty((Q,) + np.broadcast(delta, ux).shape)
    for i in range(Q):
        c = C[i]
        cu = c[0] * ux + c[1] * uy + c[2] * uz
        g[i] = W[i] * (delta + rho * (3 * cu + 4.5 * cu * cu - 1.5 * uu))
    return g


def moments(g):
    """(delta, [ux, uy, uz]) of shifted populations (lattice.js:163-174)."""
    g = np.asarray(g, dtype=np.float64)
    delta = g.sum(axis=0)
    m = [np.tensordot(C[:, a].astype(np.float64), g, axes=(0, 0)) for a in range(3)]
    inv = 1 / (1 + delta)
    return delta, [m[0] * inv, m[1] * inv, m[2] * inv]


def selfCheck() -> bool:
    """The invariants of lattice.js:184-231."""
    def fail(m):
        raise AssertionError("lattice: " + m)

    if abs(W.sum() - 1) > 1e-15:
        fail("sum(w) != 1")
    for a in range(3):
        if abs((W * C[:, a]).sum()) > 1e-15:
            fail("sum(w c) != 0")
    for a in range(3):
        for b in range(3):
            s = (W * C[:, a] * C[:, b]).sum()
            if abs(s - (CS2 if a == b else 0)) > 1e-15:
                fail("sum(w c c) != cs2 delta")
    for i in range(Q):
        if OPP[OPP[i]] != i or (C[OPP[i]] != -C[i]).any() or W[OPP[i]] != W[i]:
            fail(f"opp broken at {i}")
    if (equilibrium(0.0, [0.0, 0.0, 0.0]) != 0).any():
        fail("rest equilibrium != 0")
    d, u = 1e-4, [0.05, -0.02, 0.01]
    dd, uu = moments(equilibrium(d, u))
    if abs(dd - d) > 1e-15 or any(abs(uu[a] - u[a]) > 1e-15 for a in range(3)):
        fail("moment round trip")
    return True
