"""GPU voxelizer with the signature of voxelizar() (src/geom/voxel.js:376-433)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import VoxStats, check, lib, ptr


def voxelizar(positions, indices, g, banda: float = 1.5, selarFrestas: bool = True,
              comSdf: bool = True) -> dict:
    """Shell + flood + skin on the GPU.

    positions: float32 [nv*3] in lattice space; indices: uint32 [nt*3];
    g = {nx, ny, nz}.  Returns {'tipo': uint8[N], 'sdf': float32[N] | None,
    'stats': {...}} with the reference's field names."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    if pos.size % 3 or idx.size % 3:
        raise ValueError("positions and indices must hold whole vec3 / triangles")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    N = nx * ny * nz
    tipo = np.empty(N, dtype=np.uint8)
    sdf = np.empty(N, dtype=np.float32) if comSdf else None
    st = VoxStats()
    check(lib.cfd_voxelize(ptr(pos) if pos.size else None, pos.size // 3,
                           ptr(idx) if idx.size else None, idx.size // 3, nx, ny, nz,
                           float(banda), 1 if selarFrestas else 0, ptr(tipo),
                           ptr(sdf) if sdf is not None else None, C.byref(st)))
    stats = {
        "nCelulas": st.n_cells, "nCasca": st.n_casca, "nDentro": st.n_dentro,
        "nSolido": st.n_solido, "fracaoSolida": st.n_solido / N,
        "trianguloFora": st.triangles_outside, "ocoSuspeito": bool(st.hollow_suspect),
        "iteracoesInundacao": st.flood_iterations,
        "msRasterizacao": st.ms_raster, "msInundacao": st.ms_flood, "msTotal": st.ms_total,
    }
    return {"tipo": tipo, "sdf": sdf, "stats": stats}
