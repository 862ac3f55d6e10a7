"""dynamap_amd: MI355X (gfx950) implementation of DynaMap's dense-mapping path.

The host-side classes mirror the reference's C++ interfaces
(include/geometry/tsdf/tsdf.h:37-81 tsdfVolume, include/builder/builder.h
builder, include/solver/gnSolver/gnSolver.h:57-100 gnSolver) and call the HIP
kernels of libdynamap_hip.so through its C ABI.
"""
from . import _native
from .mapping import TsdfVolume, Session, Server
from . import configs

__all__ = ["TsdfVolume", "Session", "Server", "configs", "_native"]
