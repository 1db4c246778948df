"""Host-side solver object with the reference's API over the C ABI.

``SolverMI355X`` keeps the method names, argument meaning and error
behaviour of ``SolverWebGPU`` (src/backend/webgpu/solver.js), so code written
against the reference (and the reference's test pages, restated in tests/)
reads the same.  Async methods of the reference (lerMacros, medirForcas,
calibrarRepouso) are synchronous here: they return after a stream sync.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Sequence

import numpy as np

from . import _lib
from ._lib import CFD_STORE_FP16, CFD_STORE_FP32, Params, check, lib, ptr
from .lattice import MAGIC, Q, equilibrium


class SolverMI355X:
    """One tunnel on the GPU: populations A/B, cell types, macro field."""

    def __init__(self, handle: C.c_void_p, nx: int, ny: int, nz: int, precisao: str):
        self._h = handle
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny * nz
        self.precisao = precisao
        self.passos = 0
        self.cfg = None
        self.offsetForca = None
        self._rampa = None

    # ------------------------------------------------------------ creation
    @staticmethod
    def criar(grade, precisao: str = "fp32") -> "SolverMI355X":
        """Allocate a solver for grade = {nx, ny, nz} (solver.js:93-124).

        precisao: 'fp32' (152 B/cell/step), 'fp32exact' (the same with every
        operation rounded as the oracle rounds it: correctly rounded division
        and square root), 'fp16' (76 B/cell/step, rounded to nearest) or
        'fp16sr' (76 B/cell/step, stochastic rounding)."""
        nx, ny, nz = (int(grade["nx"]), int(grade["ny"]), int(grade["nz"])) \
            if isinstance(grade, dict) else (int(grade[0]), int(grade[1]), int(grade[2]))
        flags = {"fp32": CFD_STORE_FP32, "fp16": CFD_STORE_FP16,
                 "fp32exact": CFD_STORE_FP32 | _lib.CFD_ROUND_EXACT,
                 "fp16sr": CFD_STORE_FP16 | _lib.CFD_ROUND_STOCHASTIC}.get(precisao)
        if flags is None:
            raise ValueError("precisao must be 'fp32', 'fp32exact', 'fp16' or 'fp16sr', "
                             f"not {precisao!r}")
        h = C.c_void_p()
        check(lib.cfd_create(nx, ny, nz, flags, C.byref(h)))
        return SolverMI355X(h, nx, ny, nz, precisao)

    @property
    def bytesTotais(self) -> int:
        b = C.c_int64()
        check(lib.cfd_bytes(self._h, C.byref(b)))
        return b.value

    # ------------------------------------------------------------ configuration
    def configurar(self, omegaPlus: float, magic: float = MAGIC["WALL"], lesCs: float = 0.1,
                   inletU: Sequence[float] = (0.0, 0.0, 0.0),
                   beltU: Sequence[float] = (0.0, 0.0, 0.0), esponja=None,
                   espessuraCL: float = 0.0) -> None:
        """solver.js:277-293; esponja defaults to start round(0.82 nx), length
        max(4, round(0.18 nx))."""
        p, esp = params_struct(self.nx, omegaPlus, magic, lesCs, inletU, beltU, esponja,
                               espessuraCL)
        check(lib.cfd_configure(self._h, C.byref(p)))
        self.cfg = dict(omegaPlus=omegaPlus, magic=magic, lesCs=lesCs, inletU=list(inletU),
                        beltU=list(beltU), esponja=esp, espessuraCL=espessuraCL)

    def definirTipos(self, tipos) -> None:
        """solver.js:296-301: Uint32 cell types, N of them."""
        t = np.ascontiguousarray(tipos, dtype=np.uint32).reshape(-1)
        if t.size != self.N:
            raise ValueError(f"type field with {t.size}, expected {self.N}")
        check(lib.cfd_set_types(self._h, ptr(t), t.size))
        self.tipo = t.copy()

    def salvarCampo(self, caminho, uIn: float, formato: str = "vtk") -> None:
        """The current macro field (lerMacros) to disk: VTK with velocity,
        delta, Cp and cell type, or raw float32 [N][4] + JSON."""
        from . import campos
        m = self.lerMacros()
        g = {"nx": self.nx, "ny": self.ny, "nz": self.nz}
        if formato == "vtk":
            campos.salvar_vtk(caminho, m, g, uIn, getattr(self, "tipo", None))
        elif formato == "raw":
            campos.salvar_raw(caminho, m, g, passos=self.passos, uIn=uIn)
        else:
            raise ValueError(f"unknown format {formato!r} (vtk or raw)")

    # ------------------------------------------------------------ ramp
    def rampa(self, passos: int, aplicarVelocidade: Callable[[float], None]) -> None:
        """solver.js:323-325"""
        self._rampa = {"passos": passos, "aplicar": aplicarVelocidade, "atual": 0}

    @property
    def fatorRampa(self) -> float:
        """smoothstep t^2 (3 - 2t) of the ramp progress (solver.js:328-335)."""
        r = self._rampa
        if not r or r["atual"] >= r["passos"]:
            return 1.0
        t = r["atual"] / r["passos"]
        return t * t * (3 - 2 * t)

    # ------------------------------------------------------------ state
    def inicializar(self) -> None:
        """Free-stream equilibrium into B; front = B (solver.js:338-349)."""
        check(lib.cfd_init_equilibrium(self._h))
        self.passos = 0
        if self._rampa:
            self._rampa["atual"] = 0

    @property
    def frente(self) -> str:
        f = C.c_int()
        check(lib.cfd_get_front(self._h, C.byref(f)))
        return "B" if f.value else "A"

    @frente.setter
    def frente(self, v: str) -> None:
        if v not in ("A", "B"):
            raise ValueError("frente must be 'A' or 'B'")
        check(lib.cfd_set_front(self._h, 1 if v == "B" else 0))

    def definirCampo(self, fn, equilibrium_fn=equilibrium, vetorizado: bool = False) -> None:
        """Fill the current front from fn(x, y, z) -> {'delta', 'u'} through
        the float64 equilibrium, stored as float32 (solver.js:363-389).

        vetorizado=True calls fn once with index arrays X, Y, Z of shape
        (nz, ny, nx) and expects array-valued delta / u."""
        nx, ny, nz = self.nx, self.ny, self.nz
        if vetorizado:
            Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            r = fn(X, Y, Z) or {}
            delta = np.broadcast_to(np.asarray(r.get("delta", 0.0), dtype=np.float64), X.shape)
            u = r.get("u", (0.0, 0.0, 0.0))
            u = [np.broadcast_to(np.asarray(c, dtype=np.float64), X.shape) for c in u]
            g = equilibrium_fn(delta, u)
            dados = np.ascontiguousarray(g.reshape(Q, -1), dtype=np.float32)
        else:
            dados = np.empty((Q, self.N), dtype=np.float32)
            gb = np.empty(Q)
            for z in range(nz):
                for y in range(ny):
                    for x in range(nx):
                        cell = (z * ny + y) * nx + x
                        r = fn(x, y, z) or {}
                        equilibrium_fn(r.get("delta", 0.0), r.get("u", (0.0, 0.0, 0.0)), gb)
                        dados[:, cell] = gb
        self.definirPopulacoes(dados)
        self.passos = 0

    def definirPopulacoes(self, pops: np.ndarray) -> None:
        """Upload a (19, N) float32 population field into the current front."""
        a = np.ascontiguousarray(pops, dtype=np.float32).reshape(-1)
        check(lib.cfd_set_populations(self._h, ptr(a), a.size))

    def lerPopulacoes(self) -> np.ndarray:
        """Read back the current front as a (19, N) float32 array."""
        a = np.empty(Q * self.N, dtype=np.float32)
        check(lib.cfd_get_populations(self._h, ptr(a), a.size))
        return a.reshape(Q, self.N)

    # ------------------------------------------------------------ the loop
    def _avancar_rampa(self, n: int) -> None:
        # the ramp advances before dispatch; the caller's callback
        # reconfigures with the new factor (solver.js:396-402)
        r = self._rampa
        if r and r["atual"] < r["passos"]:
            r["atual"] += n
            if r["aplicar"]:
                r["aplicar"](self.fatorRampa)

    def passo(self, n: int = 1) -> None:
        """Enqueue n steps (solver.js:395-415); the ramp callback runs first."""
        self._avancar_rampa(n)
        check(lib.cfd_step(self._h, int(n)))
        self.passos += n

    def passoCronometrado(self, n: int) -> float:
        """passo(n) bracketed by HIP events on the solver's stream; waits for
        the n launches and returns their device time in ms."""
        self._avancar_rampa(n)
        ms = C.c_double()
        check(lib.cfd_step_timed(self._h, int(n), C.byref(ms)))
        self.passos += n
        return ms.value

    def autoajustarPasso(self, passos: int = 200, fracoes=(0.02, 0.1, 0.36),
                         blocos=(256, 512), bloco: int = 100) -> dict:
        """Launch-order autotuner (no reference counterpart): for each
        workgroup size, times `passos` steps in ascending and in alternating
        cell order and with the step start rotated by each fraction of the
        launch grid, keeps the fastest and returns the tuning.

        Steps are taken in calls of `bloco` steps (one settling call, then
        ceil(passos / bloco) timed ones per candidate), as a caller's own loop
        of passo(bloco) takes them: the inflow ramp advances and its callback
        fires at the same steps, and every order and shape updates every cell
        from the same previous state, so the run is bit-identical to one that
        calls passo(bloco) throughout.  Total:
        len(blocos) * (2 + len(fracoes)) * (1 + ceil(passos / bloco)) * bloco."""
        remap = self.ajuste["remap"]
        cand = []
        for block in blocos:
            nb = -(-self.N // block)
            span = 8 * remap if remap > 1 else 1
            grade = -(-nb // span) * span
            cand += [{"block": block, "sweep": 0, "rotate": 0},
                     {"block": block, "sweep": 1, "rotate": 0}]
            cand += [{"block": block, "sweep": 0, "rotate": max(1, int(f * grade))}
                     for f in fracoes]
        vezes = max(1, -(-passos // bloco))
        best, best_ms = None, float("inf")
        for c in cand:
            self.ajuste = c
            self.passo(bloco)          # leave the previous order's cache state behind
            ms = sum(self.passoCronometrado(bloco) for _ in range(vezes))
            if ms < best_ms:
                best, best_ms = c, ms
        self.ajuste = best
        return self.ajuste

    @staticmethod
    def passosAutoajuste(passos: int = 200, fracoes=(0.02, 0.1, 0.36), blocos=(256, 512),
                         bloco: int = 100) -> int:
        """Steps autoajustarPasso takes with these arguments."""
        return len(blocos) * (2 + len(fracoes)) * (1 + max(1, -(-passos // bloco))) * bloco

    @property
    def ajuste(self) -> dict:
        """Launch tuning and population layout of the step kernel
        (block, remap, layout 0 = SoA / 1 = 512-cell blocks, slice_pad,
        sweep 0 = ascending / 1 = alternating cell order per step, rotate R:
        with sweep 0, step k starts k*R workgroups before step 0)."""
        t = _lib.Tuning()
        check(lib.cfd_get_tuning(self._h, C.byref(t)))
        return {"block": t.block, "remap": t.remap, "layout": t.layout,
                "slice_pad": t.slice_pad, "sweep": t.sweep, "rotate": t.rotate}

    @ajuste.setter
    def ajuste(self, v: dict) -> None:
        cur = self.ajuste
        cur.update(v)
        t = _lib.Tuning(int(cur["block"]), int(cur["remap"]), int(cur["layout"]),
                        int(cur["slice_pad"]), int(cur["sweep"]), int(cur["rotate"]))
        check(lib.cfd_set_tuning(self._h, C.byref(t)))

    def sincronizar(self) -> None:
        check(lib.cfd_synchronize(self._h))

    # ------------------------------------------------------------ read-back
    def lerMacros(self) -> np.ndarray:
        """Float32 N*4 (ux, uy, uz, delta) (solver.js:427-442)."""
        m = np.empty(4 * self.N, dtype=np.float32)
        check(lib.cfd_read_macros(self._h, ptr(m), m.size))
        return m

    def atualizarMacros(self) -> None:
        """solver.js:511-520"""
        check(lib.cfd_update_macros(self._h))

    # ------------------------------------------------------------ closure
    def definirFechamento(self, fech) -> None:
        """Learned eddy viscosity in every following step (None: Smagorinsky)."""
        check(lib.cfd_set_closure(self._h, C.byref(fech.c_struct()) if fech is not None
                                  else None))
        self.fechamento = fech

    def definirDistanciaParede(self, d: np.ndarray) -> None:
        a = np.ascontiguousarray(d, dtype=np.float32).reshape(-1)
        check(lib.cfd_set_wall_distance(self._h, ptr(a), a.size))

    def calcularDistanciaParede(self) -> None:
        """Wall distance of the closure computed on the device from the types
        last set (exact Euclidean distance to the nearest wall cell, in the
        same 1/8-cell quantization as definirDistanciaParede)."""
        check(lib.cfd_compute_wall_distance(self._h))

    def lerDistanciaParede(self) -> np.ndarray:
        """The quantized wall distance the closure reads (uint8, 1/8 cell)."""
        q = np.empty(self.N, dtype=np.uint8)
        check(lib.cfd_get_wall_distance(self._h, ptr(q), q.size))
        return q

    def fallbacksFechamento(self, zerar: bool = True) -> int:
        """Fluid cell-steps that kept the Smagorinsky tau since the last reset."""
        n = C.c_int64()
        check(lib.cfd_closure_fallbacks(self._h, C.byref(n), 1 if zerar else 0))
        return n.value

    def forcaBruta(self) -> np.ndarray:
        f = (C.c_double * 3)()
        check(lib.cfd_forces(self._h, f))
        return np.array(f[:], dtype=np.float64)

    def calibrarRepouso(self) -> list:
        """Rest offset on the zeroed buffers (solver.js:480-484)."""
        self.offsetForca = [0.0, 0.0, 0.0]
        self.offsetForca = self.medirForcas()
        return self.offsetForca

    def medirForcas(self) -> list:
        """Momentum-exchange force minus the rest offset (solver.js:486-508)."""
        v = self.forcaBruta()
        o = self.offsetForca or [0.0, 0.0, 0.0]
        return [float(v[0] - o[0]), float(v[1] - o[1]), float(v[2] - o[2])]

    def massa(self) -> tuple[float, int]:
        """(float64 sum of delta over non-wall cells, non-finite cell count)."""
        m, n = C.c_double(), C.c_int64()
        check(lib.cfd_mass(self._h, C.byref(m), C.byref(n)))
        return m.value, n.value

    def destruir(self) -> None:
        """solver.js:523-527"""
        if self._h:
            lib.cfd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destruir()
        except Exception:
            pass


def params_struct(nx, omegaPlus, magic=MAGIC["WALL"], lesCs=0.1, inletU=(0.0, 0.0, 0.0),
                  beltU=(0.0, 0.0, 0.0), esponja=None, espessuraCL=0.0):
    """cfd_params of configurar(); the sponge defaults as solver.js:281-284."""
    esp = esponja if esponja is not None else {
        "inicio": _js_round(nx * 0.82),
        "comprimento": max(4, _js_round(nx * 0.18)),
    }
    p = Params()
    p.omega_plus = omegaPlus
    p.magic = magic
    p.les_cs = lesCs
    for a in range(3):
        p.belt[a] = beltU[a]
        p.inlet[a] = inletU[a]
    p.sponge_start = esp["inicio"]
    p.sponge_len = esp["comprimento"]
    p.bl_thickness = espessuraCL
    return p, esp


def _js_round(x: float) -> int:
    """Math.round: halves round toward +infinity."""
    return int(np.floor(x + 0.5))
