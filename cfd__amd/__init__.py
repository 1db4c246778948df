"""MI355X-native D3Q19 TRT-LES wind-tunnel solver and GPU voxelizer.

The compute path is libcfdmi.so (hand-written HIP for gfx950, C ABI in
include/cfd_mi355x.h).  Host helpers mirror the reference's modules:

  solver.SolverMI355X     src/backend/webgpu/solver.js (SolverWebGPU)
  voxel.voxelizar         src/geom/voxel.js
  tunel                   src/core/tunel.js
  forces                  src/core/forces.js
  units                   src/core/units.js
  validacao               src/core/validacao.js
  lattice                 src/core/lattice.js
"""
from . import forces, lattice, meshes, tunel, units, validacao  # noqa: F401

__all__ = ["forces", "lattice", "meshes", "tunel", "units", "validacao",
           "SolverMI355X", "voxelizar"]


def __getattr__(name):
    # The native pieces load libcfdmi.so, which raises when it is not built.
    if name == "SolverMI355X":
        from .solver import SolverMI355X
        return SolverMI355X
    if name == "voxelizar":
        from .voxel import voxelizar
        return voxelizar
    raise AttributeError(name)
