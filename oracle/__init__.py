"""ctypes loader of the CPU oracle (liboracle.so, built from lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker; never by the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    P = C.c_void_p
    for suf in ("f32", "f64"):
        getattr(lib, f"orc_step_{suf}").argtypes = [C.c_int] * 3 + [P, P, P, P, P]
        getattr(lib, f"orc_init_{suf}").argtypes = [C.c_int] * 3 + [P, P, P]
        getattr(lib, f"orc_macros_{suf}").argtypes = [C.c_int] * 3 + [P, P]
        getattr(lib, f"orc_forces_{suf}").argtypes = [C.c_int] * 3 + [P, P, P]
        for f in ("step", "init", "macros", "forces"):
            getattr(lib, f"orc_{f}_{suf}").restype = None
    lib.orc_step_closure_f32.argtypes = [C.c_int] * 3 + [P, P, P, P, P, P, P]
    lib.orc_step_closure_f32.restype = None
    lib.orc_closure_eval.argtypes = [P, P, C.c_size_t, P]
    lib.orc_closure_eval.restype = None
    lib.orc_voxelize.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_int, P, P, P]
    lib.orc_voxelize.restype = C.c_int
    lib.orc_equilibrium.argtypes = [C.c_double, P, P]
    lib.orc_moments.argtypes = [P, P]
    lib.orc_threads.restype = C.c_int
    lib.orc_setup()
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def params_array(omegaPlus, magic=3 / 16, lesCs=0.1, inletU=(0, 0, 0), beltU=(0, 0, 0),
                 esponja=None, espessuraCL=0.0, nx=None):
    """The 12 doubles the oracle reads: omega, magic, Cs, belt[3], inlet[3],
    sponge start, sponge length, boundary-layer thickness."""
    if esponja is None:
        esponja = {"inicio": int(np.floor(nx * 0.82 + 0.5)),
                   "comprimento": max(4, int(np.floor(nx * 0.18 + 0.5)))}
    return np.array([omegaPlus, magic, lesCs, *beltU, *inletU, esponja["inicio"],
                     esponja["comprimento"], espessuraCL], dtype=np.float64)


def step(grid, src, tipo, params, dtype=np.float32, macros=False):
    """One oracle step; returns (dst, macros or None).  Solid cells of dst keep
    the values of `src` (the reference leaves them stale; callers compare
    non-wall cells only)."""
    nx, ny, nz = grid
    src = np.ascontiguousarray(src, dtype=dtype)
    dst = src.copy()
    t = np.ascontiguousarray(tipo, dtype=np.uint32)
    m = np.zeros(4 * nx * ny * nz, dtype=dtype) if macros else None
    fn = lib().orc_step_f32 if dtype == np.float32 else lib().orc_step_f64
    fn(nx, ny, nz, _p(src), _p(dst), _p(t), _p(np.ascontiguousarray(params, np.float64)),
       _p(m) if m is not None else None)
    return dst, m


def run(grid, src, tipo, params, nsteps, dtype=np.float32):
    """nsteps A/B steps; returns (populations, macros of the last step).
    `src` is not modified."""
    a = np.array(src, dtype=dtype, copy=True, order="C")
    b = a.copy()
    nx, ny, nz = grid
    t = np.ascontiguousarray(tipo, dtype=np.uint32)
    par = np.ascontiguousarray(params, np.float64)
    m = np.zeros(4 * nx * ny * nz, dtype=dtype)
    fn = lib().orc_step_f32 if dtype == np.float32 else lib().orc_step_f64
    for k in range(nsteps):
        fn(nx, ny, nz, _p(a), _p(b), _p(t), _p(par), _p(m) if k == nsteps - 1 else None)
        a, b = b, a
    return a, m


def init(grid, tipo, params, dtype=np.float32):
    nx, ny, nz = grid
    dst = np.zeros(19 * nx * ny * nz, dtype=dtype)
    t = np.ascontiguousarray(tipo, dtype=np.uint32)
    fn = lib().orc_init_f32 if dtype == np.float32 else lib().orc_init_f64
    fn(nx, ny, nz, _p(dst), _p(t), _p(np.ascontiguousarray(params, np.float64)))
    return dst


def macros(grid, pops, dtype=np.float32):
    nx, ny, nz = grid
    p = np.ascontiguousarray(pops, dtype=dtype)
    m = np.zeros(4 * nx * ny * nz, dtype=dtype)
    fn = lib().orc_macros_f32 if dtype == np.float32 else lib().orc_macros_f64
    fn(nx, ny, nz, _p(p), _p(m))
    return m


def forces(grid, pops, tipo, dtype=np.float32):
    nx, ny, nz = grid
    p = np.ascontiguousarray(pops, dtype=dtype)
    t = np.ascontiguousarray(tipo, dtype=np.uint32)
    out = np.zeros(3, dtype=np.float64)
    fn = lib().orc_forces_f32 if dtype == np.float32 else lib().orc_forces_f64
    fn(nx, ny, nz, _p(p), _p(t), _p(out))
    return out


def voxelize(pos, idx, grid, banda=1.5, selar=True):
    """Returns (tipo uint8[N], sdf float32[N], stats int64[5])."""
    nx, ny, nz = grid
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    N = nx * ny * nz
    tipo = np.empty(N, dtype=np.uint8)
    sdf = np.empty(N, dtype=np.float32)
    st = np.zeros(5, dtype=np.int64)
    rc = lib().orc_voxelize(_p(pos) if pos.size else None, pos.size // 3,
                            _p(idx) if idx.size else None, idx.size // 3, nx, ny, nz,
                            float(banda), 1 if selar else 0, _p(tipo), _p(sdf), _p(st))
    if rc != 0:
        raise MemoryError("oracle voxelizer allocation failed")
    return tipo, sdf, st


def threads() -> int:
    return lib().orc_threads()


def closure_array(params, shift, scale, lo, hi, out_scale, ativacao="silu"):
    """The float32 block orc_closure reads: 1377 MLP parameters (cfd_mlp
    order), then shift, scale, lo, hi (8 each), out_scale and the hidden
    activation (0 SiLU, 1 hard-SiLU)."""
    codigo = {"silu": 0.0, "hard_silu": 1.0}[ativacao]
    return np.ascontiguousarray(np.concatenate([
        np.asarray(params, np.float32).reshape(-1), np.asarray(shift, np.float32),
        np.asarray(scale, np.float32), np.asarray(lo, np.float32), np.asarray(hi, np.float32),
        np.array([out_scale, codigo], np.float32)]), dtype=np.float32)


def closure_eval(cl, X):
    """nu of the learned closure for raw features X (n, 8); NaN outside the range."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.empty(X.shape[0], dtype=np.float32)
    lib().orc_closure_eval(_p(cl), _p(X), X.shape[0], _p(out))
    return out


def step_closure(grid, src, tipo, params, cl, wdist, macros=False):
    """One float32 oracle step with the learned closure (cl from
    closure_array, wdist: float32 per cell)."""
    nx, ny, nz = grid
    src = np.ascontiguousarray(src, dtype=np.float32)
    dst = src.copy()
    t = np.ascontiguousarray(tipo, dtype=np.uint32)
    m = np.zeros(4 * nx * ny * nz, dtype=np.float32) if macros else None
    wd = np.ascontiguousarray(wdist, dtype=np.float32)
    lib().orc_step_closure_f32(nx, ny, nz, _p(src.reshape(-1)), _p(dst.reshape(-1)), _p(t),
                               _p(np.ascontiguousarray(params, dtype=np.float64)),
                               _p(m) if m is not None else None, _p(cl), _p(wd))
    return dst, m
