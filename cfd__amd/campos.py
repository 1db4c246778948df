"""Field output: the macro field the renderers consume, written to disk.

The solver's macro field is float32 [N][4] = (ux, uy, uz, delta) in lattice
units, zero in walls (wgsl.js:255; lerMacros).  The pressure coefficient the
reference's Cp map draws is Cp = c_s^2 delta / (1/2 u_in^2) with c_s^2 = 1/3
(renderer.js:419-444).  Two formats:

* raw: the float32 [N][4] array as it comes from lerMacros, little-endian,
  cell = (z * ny + y) * nx + x, plus a JSON side file with the grid;
* VTK legacy binary STRUCTURED_POINTS (ParaView, VisIt): velocity vectors,
  delta, Cp and the cell type as point data.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

CS2 = 1.0 / 3.0


def cp(macros: np.ndarray, u_in: float) -> np.ndarray:
    """Pressure coefficient per cell from the macro field."""
    m = np.asarray(macros, dtype=np.float32).reshape(-1, 4)
    return (CS2 * m[:, 3].astype(np.float64) / (0.5 * u_in * u_in)).astype(np.float32)


def salvar_raw(caminho, macros: np.ndarray, grade: dict, **meta) -> None:
    """macros.raw (float32 [N][4]) + macros.json (grid and metadata)."""
    caminho = Path(caminho)
    m = np.ascontiguousarray(macros, dtype="<f4").reshape(-1)
    N = grade["nx"] * grade["ny"] * grade["nz"]
    if m.size != 4 * N:
        raise ValueError(f"macro field with {m.size} values, expected {4 * N}")
    caminho.write_bytes(m.tobytes())
    caminho.with_suffix(".json").write_text(json.dumps(
        {"nx": grade["nx"], "ny": grade["ny"], "nz": grade["nz"], "layout": "[z][y][x][4]",
         "componentes": ["ux", "uy", "uz", "delta"], "dtype": "float32-le", **meta}))


def ler_raw(caminho) -> tuple[np.ndarray, dict]:
    caminho = Path(caminho)
    meta = json.loads(caminho.with_suffix(".json").read_text())
    m = np.frombuffer(caminho.read_bytes(), dtype="<f4").reshape(-1, 4)
    return m, meta


def salvar_vtk(caminho, macros: np.ndarray, grade: dict, u_in: float,
               tipo: np.ndarray | None = None, titulo: str = "cfd__amd macros") -> None:
    """VTK legacy binary (big-endian) structured points with VECTORS velocity
    and SCALARS delta, Cp (and tipo when given); spacing 1 cell."""
    nx, ny, nz = grade["nx"], grade["ny"], grade["nz"]
    N = nx * ny * nz
    m = np.asarray(macros, dtype=np.float32).reshape(-1, 4)
    if m.shape[0] != N:
        raise ValueError(f"macro field with {m.shape[0]} cells, expected {N}")
    out = bytearray()
    out += (f"# vtk DataFile Version 3.0\n{titulo[:200]}\nBINARY\nDATASET STRUCTURED_POINTS\n"
            f"DIMENSIONS {nx} {ny} {nz}\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA {N}\n").encode()
    out += b"VECTORS velocity float\n" + np.ascontiguousarray(m[:, :3], dtype=">f4").tobytes() + b"\n"
    out += (b"SCALARS delta float 1\nLOOKUP_TABLE default\n" +
            np.ascontiguousarray(m[:, 3], dtype=">f4").tobytes() + b"\n")
    out += (b"SCALARS Cp float 1\nLOOKUP_TABLE default\n" +
            cp(m, u_in).astype(">f4").tobytes() + b"\n")
    if tipo is not None:
        t = np.asarray(tipo).reshape(-1)
        if t.size != N:
            raise ValueError(f"type field with {t.size} cells, expected {N}")
        out += (b"SCALARS tipo int 1\nLOOKUP_TABLE default\n" +
                t.astype(">i4").tobytes() + b"\n")
    Path(caminho).write_bytes(bytes(out))


def ler_vtk(caminho) -> dict:
    """Reads back what salvar_vtk writes (for checks and simple tools)."""
    b = Path(caminho).read_bytes()
    head_end = b.index(b"POINT_DATA")
    head = b[:head_end].decode()
    dims = [int(v) for v in head.split("DIMENSIONS")[1].split("\n")[0].split()]
    N = dims[0] * dims[1] * dims[2]
    out = {"dims": dims}
    p = b.index(b"\n", head_end) + 1
    while p < len(b):
        line_end = b.index(b"\n", p)
        words = b[p:line_end].decode().split()
        p = line_end + 1
        if words[0] == "VECTORS":
            out[words[1]] = np.frombuffer(b, dtype=">f4", count=3 * N, offset=p).reshape(N, 3)
            p += 12 * N + 1
        elif words[0] == "SCALARS":
            p = b.index(b"\n", p) + 1          # LOOKUP_TABLE line
            dt = ">f4" if words[2] == "float" else ">i4"
            out[words[1]] = np.frombuffer(b, dtype=dt, count=N, offset=p)
            p += 4 * N + 1
        else:
            break
    return out
