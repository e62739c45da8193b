"""MI355X (gfx950) implementation of REEF3D's CFD free-surface hot path.

The compute runs in hand-written HIP kernels in libreef3d_hip.so (C-ABI in
include/reef3d_hip.h); `reef3d_amd.hotpath.Domain` is the Python host mirror of the
reference's convection / reini / solver / pressure / ghostcell interfaces.
"""
from .hotpath import Batch, Domain, ReefError, Server, FIELDS, library  # noqa: F401

__all__ = ["Batch", "Domain", "ReefError", "Server", "FIELDS", "library"]
