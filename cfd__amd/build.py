"""Build the native parts of the package in-tree.

* ``cfd__amd/libcfdmi.so`` - the HIP kernels and the C ABI of
  ``include/cfd_mi355x.h``, compiled for gfx950 with hipcc.
* ``cfd__amd/node/cfdmi.node`` - the N-API addon that binds that C ABI for
  Node.js (only when the Node headers are installed).

Run ``python -m cfd__amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libcfdmi.so"
ADDON = PKG / "node" / "cfdmi.node"
NODE_INC = Path("/usr/include/node")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CFD_OFFLOAD_ARCH", "gfx950")

# Every HIP translation unit: packed-fp32 instructions are switched off (the
# compiler would otherwise form v_pk_* from ordinary float code).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# Per-file extras: the voxelizer must not contract a*b+c into fma, since it
# reproduces IEEE float64 JavaScript arithmetic bit for bit.
# The lattice kernels round every product before the following add unless the
# source fuses it explicitly (fmaf): every instantiation of the step (plain,
# pull-only, with the force accumulation or the batch loop around it, the
# closure's fallback) then rounds identically, and the exact-rounding mode
# identically to a plain IEEE float32 CPU build.
# The lattice kernels also flush float32 subnormals to zero (the WGSL the
# reference emits leaves subnormals to the implementation, which may flush
# them): no population is ever subnormal in the workloads, and the step ran
# 1-2 % faster in the first hundred steps after an initialization this way.
FTZ = ["-fgpu-flush-denormals-to-zero"]
FILE_FLAGS = {
    "voxel.hip": ["-ffp-contract=off"],
    "lbm_kernels.hip": ["-ffp-contract=off", *FTZ],
    "closure.hip": ["-ffp-contract=off", *FTZ],
}
SOURCES = ["lbm_kernels.hip", "closure.hip", "cfd_core.hip", "batch_api.hip", "voxel.hip"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_lib(verbose: bool = False) -> Path:
    """Compile every HIP source to an object, then link libcfdmi.so."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_addon(verbose: bool = False) -> Path | None:
    """Compile the N-API addon against libcfdmi.so (needs the Node headers)."""
    if not (NODE_INC / "node_api.h").exists():
        return None
    src = PKG / "node" / "addon.cc"
    if not src.exists():
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{NODE_INC}", f"-I{ROOT / 'include'}",
           str(src), "-o", str(ADDON), f"-L{PKG}", "-lcfdmi", "-Wl,-rpath,$ORIGIN/..",
           "-Wl,--allow-shlib-undefined"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return ADDON


def build_all(verbose: bool = False) -> None:
    build_lib(verbose)
    build_addon(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(LIB)
