"""dynamap_amd: MI355X (gfx950) implementation of DynaMap's dense-mapping path.

The host-side classes mirror the reference's C++ interfaces
(include/geometry/tsdf/tsdf.h:37-81 tsdfVolume, include/builder/builder.h
builder, include/solver/gnSolver/gnSolver.h:57-100 gnSolver,
include/geometry/mesh/meshExtractor/meshExtractor.h:20-59 MeshExtractor,
include/geometry/mesh/mesh.h:25-66 Mesh, include/geometry/defGraph/defGraph.h
defGraph, include/blender/dqBlender/dqBlender.h dqBlender,
include/solver/nonRigidICP/nonRigidICP.h nonRigidICP) and call the HIP
kernels of libdynamap_hip.so through its C ABI.
"""
from . import _native
from .mapping import (TsdfVolume, Session, Server, Mesh, MeshExtractor, mc_tri_table, DefGraph, DqBlender,
                      NonRigidICP, NonRigidBatch, bilateral_filter, depth_pyramid)
from . import configs

__all__ = ["TsdfVolume", "Session", "Server", "Mesh", "MeshExtractor", "mc_tri_table", "DefGraph", "DqBlender",
           "NonRigidICP", "NonRigidBatch", "bilateral_filter", "depth_pyramid", "configs", "_native"]
