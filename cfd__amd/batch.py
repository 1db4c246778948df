"""Many independent tunnels advanced by one cooperative launch.

LoteMI355X(tuneis) takes a list of {'grade': {nx, ny, nz}, 'tipo': uint32[N],
'cfg': keyword arguments of SolverMI355X.configurar}.  rodar(passos) runs
every tunnel `passos` steps inside one launch (smallest tunnels admitted
first while the admitted tiles stay within the cap); resultados() returns the
force samples (the momentum exchange every `amostra` steps, float64) and
each tunnel's completion time.  autoajustar() times short runs of candidate
tile shapes and launch sizes and keeps the fastest.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BatchTuning, TunnelSpec, check, lib, ptr
from .lattice import Q
from .solver import params_struct


class LoteMI355X:
    def __init__(self, tuneis: list[dict]):
        self.n = len(tuneis)
        self.grades = [t["grade"] for t in tuneis]
        self._tipos = [np.ascontiguousarray(t["tipo"], dtype=np.uint32).reshape(-1)
                       for t in tuneis]
        specs = (TunnelSpec * self.n)()
        for k, t in enumerate(tuneis):
            g = t["grade"]
            N = g["nx"] * g["ny"] * g["nz"]
            if self._tipos[k].size != N:
                raise ValueError(f"tunnel {k}: type field with {self._tipos[k].size}, "
                                 f"expected {N}")
            p, _ = params_struct(g["nx"], **t["cfg"])
            specs[k].nx, specs[k].ny, specs[k].nz = g["nx"], g["ny"], g["nz"]
            specs[k].tipo = ptr(self._tipos[k])
            specs[k].params = p
        self._h = C.c_void_p()
        check(lib.cfd_batch_create(specs, self.n, C.byref(self._h)))
        self.passos = 0
        self.amostra = 0
        self.ajuste = None

    @property
    def celulas(self) -> np.ndarray:
        return np.array([g["nx"] * g["ny"] * g["nz"] for g in self.grades], dtype=np.int64)

    @staticmethod
    def limite_residencia(block: int = 512) -> tuple[int, int]:
        """(workgroups per CU, CUs) of the batch kernel."""
        a, b = C.c_int(), C.c_int()
        check(lib.cfd_batch_max_grid(block, C.byref(a), C.byref(b)))
        return a.value, b.value

    def reiniciar(self) -> None:
        check(lib.cfd_batch_reset(self._h))

    def rodar(self, passos: int, amostra: int = 100, ajuste: dict | None = None) -> float:
        """One launch of `passos` steps per tunnel; returns its device time (ms)."""
        t = None
        if ajuste is not None:
            t = BatchTuning(int(ajuste["block"]), int(ajuste["tile_cells"]),
                            int(ajuste.get("wg_per_cu", 0)),
                            int(ajuste.get("max_active_tiles", 0)),
                            int(ajuste.get("tiles_per_grab", 0)))
        ms = C.c_double()
        check(lib.cfd_batch_run(self._h, int(passos), int(amostra),
                                C.byref(t) if t is not None else None, C.byref(ms)))
        self.passos, self.amostra = passos, amostra
        if ajuste is not None:
            self.ajuste = dict(ajuste)
        return ms.value

    def resultados(self) -> dict:
        ns, rounds = C.c_int(), C.c_int()
        check(lib.cfd_batch_results(self._h, None, None, C.byref(ns), C.byref(rounds)))
        f = np.zeros((self.n, max(ns.value, 1), 3), dtype=np.float64)
        fin = np.zeros(self.n, dtype=np.float64)
        check(lib.cfd_batch_results(self._h, ptr(f), ptr(fin), C.byref(ns), C.byref(rounds)))
        return {"forcas": f[:, :ns.value], "fim_ms": fin, "rodadas": rounds.value}

    def populacoes(self, k: int) -> np.ndarray:
        g = self.grades[k]
        N = g["nx"] * g["ny"] * g["nz"]
        a = np.empty(Q * N, dtype=np.float32)
        check(lib.cfd_batch_get_populations(self._h, k, ptr(a), a.size))
        return a.reshape(Q, N)

    def definirPopulacoes(self, k: int, pops: np.ndarray) -> None:
        a = np.ascontiguousarray(pops, dtype=np.float32).reshape(-1)
        check(lib.cfd_batch_set_populations(self._h, k, ptr(a), a.size))

    def autoajustar(self, passos: int = 40, candidatos: list[dict] | None = None) -> dict:
        """Times `passos` steps per candidate (state restored afterwards) and
        keeps the fastest; returns {'melhor': ..., 'medidas': [...]}."""
        if candidatos is None:
            # tile shape x workgroup size x resident workgroups per CU (the
            # most that fit, and one fewer: fewer tiles in flight thrash the
            # L2 less) x admission cap; the cap is given in cells (mult x 4096
            # per workgroup that fits) so that every tile size and residency
            # admits the same amount of work per round
            candidatos = []
            for block in (256, 512):
                pmax, cus = self.limite_residencia(block)
                for per in sorted({max(1, pmax - 1), pmax}):
                    for tile in (512, 1024, 2048, 4096):
                        for mult in (8, 32):
                            candidatos.append({"block": block, "tile_cells": tile,
                                               "wg_per_cu": per, "tiles_per_grab": 1,
                                               "max_active_tiles": pmax * cus * mult * 4096 // tile})
        medidas = []
        for c in candidatos:
            ms = self.rodar(passos, 0, c)
            medidas.append((ms, c))
        self.reiniciar()
        melhor = min(medidas, key=lambda m: m[0])[1]
        self.ajuste = dict(melhor)
        return {"melhor": melhor, "medidas": [{"ms": m, **c} for m, c in medidas]}

    def destruir(self) -> None:
        if self._h:
            lib.cfd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destruir()
        except Exception:
            pass
