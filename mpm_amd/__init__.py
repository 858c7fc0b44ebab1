"""MI355X-native explicit MPM snow solver (gfx950 HIP kernels behind a C-ABI).

Python side: ctypes binding (``native``), the reference-shaped solver
classes (``solver``), seeded workloads (``scenes``), batching and serving
(``batch``, ``serve``).  The compute path is ``lib/libmpm_hip.so``; there is
no CPU fallback.
"""
from . import native
from .native import (APIC, CUBIC_BSPLINE, FIXED_COROTATED, LINEAR, NEO_HOOKEAN, PIC_FLIP, QUADRATIC_BSPLINE,
                     MpmError)
from .solver import (APICMPMSolver3D, MPMSolver3D, ParticleInput, ParticleState, Solver, SolverDesc,
                     debug_svd3)
from .batch import Batch, Tune, tune_lookup

__all__ = [
    "native", "Solver", "SolverDesc", "ParticleInput", "ParticleState", "MPMSolver3D", "APICMPMSolver3D",
    "debug_svd3", "MpmError", "LINEAR", "QUADRATIC_BSPLINE", "CUBIC_BSPLINE", "FIXED_COROTATED", "NEO_HOOKEAN",
    "PIC_FLIP", "APIC", "Batch", "Tune", "tune_lookup",
]
