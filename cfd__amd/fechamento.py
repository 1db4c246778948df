"""Learned sub-grid closure: the MLP, its training and its use in the solver.

The closure replaces the Smagorinsky eddy viscosity of the step
(ir.js:262-301) by nu_t = out_scale * softplus(MLP(x_n)), with x_n the
normalized features (Pi_xx, Pi_yy, Pi_zz, Pi_xy, Pi_xz, Pi_yz) / rho of the
non-equilibrium stress, |S| = 3 Q_s / (2 rho tau_s) and the wall distance in
cells.  The hidden activation is SiLU or hard-SiLU, z * clamp(z/6 + 1/2, 0, 1)
(``ativacao``); hard-SiLU costs the step's matrix-core path one clamped fma
and one multiply per hidden unit instead of an exponential and a reciprocal.
Training (cfd_mlp_train) and inference (inside the step, bf16 on the
matrix cores) run on the GPU through the C ABI; this module prepares data and
parameters.

    fech = Fechamento.novo(X_raw, y_nu)        # normalization from the data
    fech.treinar(X_raw, y_nu, passos=20000)    # Adam, lr 1e-3
    solver.definirFechamento(fech)
    solver.calcularDistanciaParede()           # on the device, from the types

distancia_parede (scipy, host) gives the same distances unquantized; it is
what the training features use.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import CFD_ACT_HARD_SILU, CFD_ACT_SILU, Closure, Mlp, TrainCfg, check, lib, ptr

N_PARAMS = 1377
FEATURES = ("Pi_xx", "Pi_yy", "Pi_zz", "Pi_xy", "Pi_xz", "Pi_yz", "|S|", "wall_distance")
OFF = {"w1": 0, "b1": 256, "w2": 288, "b2": 1312, "w3": 1344, "b3": 1376}
ATIVACOES = {"silu": CFD_ACT_SILU, "hard_silu": CFD_ACT_HARD_SILU}


def _codigo(ativacao: str) -> int:
    if ativacao not in ATIVACOES:
        raise ValueError(f"unknown activation {ativacao!r} (one of {sorted(ATIVACOES)})")
    return ATIVACOES[ativacao]


def iniciar_mlp(seed: int = 0, bias_saida: float = -2.0) -> np.ndarray:
    """He-uniform weights, zero hidden biases; the output bias starts the
    softplus in its low range (small initial eddy viscosity)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    p[OFF["w1"]:OFF["b1"]] = rng.uniform(-1, 1, 256) * np.sqrt(6 / 8)
    p[OFF["w2"]:OFF["b2"]] = rng.uniform(-1, 1, 1024) * np.sqrt(6 / 32)
    p[OFF["w3"]:OFF["b3"]] = rng.uniform(-1, 1, 32) * np.sqrt(6 / 32) * 0.1
    p[OFF["b3"]] = bias_saida
    return p


def _mlp_struct(p: np.ndarray, ativacao: str = "silu") -> Mlp:
    m = Mlp()
    C.memmove(C.addressof(m), np.ascontiguousarray(p, dtype=np.float32).ctypes.data,
              N_PARAMS * 4)
    m.activation = _codigo(ativacao)
    return m


def _mlp_array(m: Mlp) -> np.ndarray:
    out = np.empty(N_PARAMS, dtype=np.float32)
    C.memmove(out.ctypes.data, C.addressof(m), N_PARAMS * 4)
    return out


def gradiente(Xn: np.ndarray, y: np.ndarray, params: np.ndarray,
              ativacao: str = "silu") -> tuple[np.ndarray, float]:
    """Full-batch gradient of mean (softplus(MLP(Xn)) - y)^2 (float32 kernels)."""
    Xn = np.ascontiguousarray(Xn, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = Mlp()
    loss = C.c_double()
    check(lib.cfd_mlp_grad(ptr(Xn), ptr(y), y.size, C.byref(_mlp_struct(params, ativacao)),
                           C.byref(g), C.byref(loss)))
    return _mlp_array(g), loss.value


def treinar_mlp(Xn: np.ndarray, y: np.ndarray, params: np.ndarray, passos: int = 20000,
                lote: int = 4096, lr: float = 1e-3, seed: int = 1,
                beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
                ativacao: str = "silu"):
    """Adam on minibatches drawn on the device; returns (params, loss curve)."""
    Xn = np.ascontiguousarray(Xn, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    cfg = TrainCfg(int(passos), int(lote), float(lr), float(beta1), float(beta2), float(eps),
                   int(seed))
    m = _mlp_struct(params, ativacao)
    curva = np.empty(int(passos), dtype=np.float32)
    check(lib.cfd_mlp_train(ptr(Xn), ptr(y), y.size, C.byref(cfg), C.byref(m), ptr(curva)))
    return _mlp_array(m), curva


@dataclass
class Fechamento:
    params: np.ndarray
    shift: np.ndarray
    scale: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    out_scale: float
    ativacao: str = "silu"
    historico: dict = field(default_factory=dict)

    @classmethod
    def novo(cls, X: np.ndarray, y_nu: np.ndarray, seed: int = 0, margem: float = 0.5,
             ativacao: str = "silu"):
        """Normalization (mean / std per feature), the trained range (the
        data's min..max of x_n widened by `margem` standard deviations) and
        the output scale (std of the target) from the training data."""
        X = np.asarray(X, dtype=np.float64)
        mu = X.mean(axis=0)
        sd = X.std(axis=0)
        sd[sd < 1e-30] = 1.0
        xn = (X - mu) / sd
        lo = xn.min(axis=0) - margem
        hi = xn.max(axis=0) + margem
        out = float(np.std(y_nu)) or 1.0
        _codigo(ativacao)
        return cls(iniciar_mlp(seed), mu.astype(np.float32), (1 / sd).astype(np.float32),
                   lo.astype(np.float32), hi.astype(np.float32), out, ativacao)

    def normalizar(self, X: np.ndarray) -> np.ndarray:
        return ((np.asarray(X, np.float32) - self.shift) * self.scale).astype(np.float32)

    def treinar(self, X: np.ndarray, y_nu: np.ndarray, passos: int = 20000, lote: int = 4096,
                lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        self.params, curva = treinar_mlp(self.normalizar(X), np.asarray(y_nu) / self.out_scale,
                                         self.params, passos, lote, lr, seed,
                                         ativacao=self.ativacao)
        self.historico["perda"] = curva
        return curva

    def c_struct(self) -> Closure:
        c = Closure()
        c.mlp = _mlp_struct(self.params, self.ativacao)
        for k in range(8):
            c.in_shift[k] = float(self.shift[k])
            c.in_scale[k] = float(self.scale[k])
            c.in_lo[k] = float(self.lo[k])
            c.in_hi[k] = float(self.hi[k])
        c.out_scale = float(self.out_scale)
        return c

    def inferir(self, X: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """nu_t of raw features X (n, 8) through the matrix-core path and
        whether each row lies inside the trained range."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        nu = np.empty(n, dtype=np.float32)
        dentro = np.empty(n, dtype=np.uint8)
        check(lib.cfd_mlp_infer(ptr(X), n, C.byref(self.c_struct()), ptr(nu), ptr(dentro)))
        return nu, dentro.astype(bool)

    def salvar(self, caminho) -> None:
        np.savez(caminho, params=self.params, shift=self.shift, scale=self.scale, lo=self.lo,
                 hi=self.hi, out_scale=np.float32(self.out_scale),
                 ativacao=np.array(self.ativacao))

    def salvarJson(self, caminho) -> None:
        """The model as JSON for the Node class (carregarFechamento in
        solver.js): every float32 written as the shortest decimal that reads
        back to it."""
        import json

        def f32(a):
            a = np.asarray(a, dtype=np.float32).reshape(-1)
            if not np.isfinite(a).all():
                raise ValueError("salvarJson: JSON has no infinities")
            return [float(str(v)) for v in a]
        d = {"params": f32(self.params), "shift": f32(self.shift), "scale": f32(self.scale),
             "lo": f32(self.lo), "hi": f32(self.hi),
             "out_scale": f32([self.out_scale])[0], "ativacao": self.ativacao}
        with open(caminho, "w") as f:
            json.dump(d, f)

    @classmethod
    def carregar(cls, caminho) -> "Fechamento":
        d = np.load(caminho, allow_pickle=False)
        ativacao = str(d["ativacao"]) if "ativacao" in d.files else "silu"
        _codigo(ativacao)
        return cls(d["params"], d["shift"], d["scale"], d["lo"], d["hi"], float(d["out_scale"]),
                   ativacao)


def distancia_parede(tipo: np.ndarray, grade: dict, paredes=(1, 4, 7)) -> np.ndarray:
    """Euclidean distance (cells) from every cell centre to the nearest wall
    cell centre (SOLIDO, SOLIDO_MOVEL, PAREDE); 0 inside walls."""
    from scipy.ndimage import distance_transform_edt

    nx, ny, nz = grade["nx"], grade["ny"], grade["nz"]
    t = np.asarray(tipo).reshape(nz, ny, nx)
    parede = np.isin(t, paredes)
    if not parede.any():
        return np.full(nx * ny * nz, 255.0 / 8.0, dtype=np.float32)
    return distance_transform_edt(~parede).astype(np.float32).reshape(-1)
