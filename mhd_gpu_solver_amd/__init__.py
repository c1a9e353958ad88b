"""MI355X-native ``hip`` backend for the 2-D GLM-MHD finite-volume solver.

Hand-written HIP kernels (gfx950) for the SSP-RK2 MUSCL + HLL + GLM (+ Ohmic/Hall/ambipolar)
step, behind a C ABI (include/mhd_hip.h) and a Python mirror of the reference's ``mhd_solver``
module. Build the library with ``python -m mhd_gpu_solver_amd.build``.
"""
from . import state_io  # noqa: F401
from .batch import BatchEngine, CapacityFull, instance_bytes  # noqa: F401
from .solver import (  # noqa: F401
    FIELDS,
    BoundaryConditionType,
    DeviceSolver,
    MHDSolver,
    NonIdealConfig,
    SlopeLimiter,
    SolveParams,
    SolveResult,
    available_backends,
    cuda_available,
    hip_available,
    init_ic,
    metal_available,
    stream_bandwidth,
    wall_plan,
)

__all__ = [
    "BatchEngine", "CapacityFull", "instance_bytes", "FIELDS", "BoundaryConditionType", "DeviceSolver", "MHDSolver", "NonIdealConfig",
    "SlopeLimiter", "SolveParams", "SolveResult", "available_backends", "cuda_available",
    "hip_available", "init_ic", "metal_available", "stream_bandwidth", "wall_plan", "state_io",
]
