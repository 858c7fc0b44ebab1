"""ctypes binding of the C-ABI in ``include/mpm_hip.h`` (``libmpm_hip.so``).

The library is built in-tree (``python -m mpm_amd.build``).  There is no
fallback: if the shared object is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmpm_hip.so")

MPM_OK = 0
MPM_ERR_INVALID = -1
MPM_ERR_HIP = -2
MPM_ERR_OOM = -3
MPM_ERR_UNSUPPORTED = -4
MPM_ERR_STATE = -5
MPM_ERR_NOT_FOUND = -6
MPM_ERR_IO = -7

# enums (reference include/enums.h:6-15)
LINEAR, QUADRATIC_BSPLINE, CUBIC_BSPLINE = 0, 1, 2
FIXED_COROTATED, NEO_HOOKEAN = 0, 1
PIC_FLIP, APIC = 0, 1


class MpmDesc(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3),
        ("resolution", C.c_int32 * 3),
        ("stride", C.c_float),
        ("friction", C.c_float),
        ("blend", C.c_float),
        ("interp", C.c_int32),
        ("transfer", C.c_int32),
        ("dt", C.c_float),
        ("gravity", C.c_float),
        ("particle_collision", C.c_int32),
        ("reserved", C.c_int32 * 4),
    ]


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class MpmParticleInput(C.Structure):
    _fields_ = [
        ("count", C.c_int64),
        ("position", _fp),
        ("velocity", _fp),
        ("mass", _fp),
        ("hardening", _fp),
        ("young", _fp),
        ("poisson", _fp),
        ("critical_compression", _fp),
        ("critical_stretch", _fp),
        ("model", _ip),
    ]


class MpmParticleState(C.Structure):
    _fields_ = [
        ("count", C.c_int64),
        ("volume0", _fp),
        ("mass", _fp),
        ("position", _fp),
        ("velocity", _fp),
        ("Fe", _fp),
        ("Fp", _fp),
        ("mu0", _fp),
        ("lambda0", _fp),
        ("hardening", _fp),
        ("critical_compression", _fp),
        ("critical_stretch", _fp),
        ("model", _ip),
        ("B", _fp),
    ]


class MpmGridHost(C.Structure):
    _fields_ = [
        ("nodes", C.c_int64),
        ("mass", _fp),
        ("velocity", _fp),
        ("velocity_star", _fp),
        ("force", _fp),
    ]


class MpmStats(C.Structure):
    _fields_ = [
        ("particles", C.c_int64),
        ("steps", C.c_int64),
        ("active_nodes", C.c_int64),
        ("active_nodes_sum", C.c_int64),
        ("work_items", C.c_int64),
        ("escaped", C.c_int64),
        ("device_bytes", C.c_int64),
        ("last_step_ms", C.c_double),
        ("max_speed", C.c_double),
        ("cfl", C.c_double),
        ("graph_captures", C.c_int64),
        ("incremental_sorts", C.c_int64),
    ]


class MpmKernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double)]


class MpmTune(C.Structure):
    _fields_ = [
        ("item_max", C.c_int32),
        ("g2p_threads", C.c_int32),
        ("workgroups", C.c_int32),
        ("bucket", C.c_int32),
        ("step_ms", C.c_double),
        ("fuse_transfer", C.c_int32),
    ]


class MpmBatchStats(C.Structure):
    _fields_ = [
        ("scenes", C.c_int64),
        ("particles", C.c_int64),
        ("steps", C.c_int64),
        ("storage_blocks", C.c_int64),
        ("key_bits", C.c_int64),
        ("work_items", C.c_int64),
        ("device_bytes", C.c_int64),
        ("last_step_ms", C.c_double),
        ("tune", MpmTune),
        ("incremental_sorts", C.c_int64),
    ]


class MpmSceneHealth(C.Structure):
    _fields_ = [("max_speed", C.c_double), ("cfl", C.c_double), ("nonfinite", C.c_int64), ("escaped", C.c_int64)]


class MpmCacheInfo(C.Structure):
    _fields_ = [("cached_bytes", C.c_int64), ("live_bytes", C.c_int64), ("hits", C.c_int64), ("misses", C.c_int64)]


# Every exported entry point and its signature (also used by the CPU test
# that checks the library exports what the header declares).
SIGNATURES = {
    "mpm_create": (C.c_int, [C.POINTER(MpmDesc), C.POINTER(MpmParticleInput), C.POINTER(C.c_void_p)]),
    "mpm_create_state": (C.c_int, [C.POINTER(MpmDesc), C.POINTER(MpmParticleState), C.c_int, C.POINTER(C.c_void_p)]),
    "mpm_destroy": (None, [C.c_void_p]),
    "mpm_last_error": (C.c_char_p, []),
    "mpm_version": (C.c_char_p, []),
    "mpm_set_cache_limit": (C.c_int, [C.c_int64]),
    "mpm_get_cache_info": (C.c_int, [C.POINTER(MpmCacheInfo)]),
    "mpm_release_cache": (C.c_int, []),
    "mpm_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_synchronize": (C.c_int, [C.c_void_p]),
    "mpm_step_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_initialize": (C.c_int, [C.c_void_p]),
    "mpm_reset_grid": (C.c_int, [C.c_void_p]),
    "mpm_particles_to_grid": (C.c_int, [C.c_void_p]),
    "mpm_compute_external_forces": (C.c_int, [C.c_void_p]),
    "mpm_compute_gravity_forces": (C.c_int, [C.c_void_p]),
    "mpm_update_grid_velocities": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_solve_linear_system": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_grid_collision": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_grid_to_particles": (C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    "mpm_particles_collision": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_advect_particles": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_switch_particles_collision": (C.c_int, [C.c_void_p]),
    "mpm_set_particles_collision": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_set_grid_record": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_set_reorder_interval": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_set_transfer_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_sort": (C.c_int, [C.c_void_p]),
    "mpm_read_particles": (C.c_int, [C.c_void_p, C.POINTER(MpmParticleState)]),
    "mpm_read_particles_noalloc": (C.c_int, [C.c_void_p, C.POINTER(MpmParticleState)]),
    "mpm_write_particles": (C.c_int, [C.c_void_p, C.POINTER(MpmParticleState)]),
    "mpm_read_grid": (C.c_int, [C.c_void_p, C.POINTER(MpmGridHost)]),
    "mpm_read_positions4": (C.c_int, [C.c_void_p, _fp]),
    "mpm_pack_positions4": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mpm_positions4_async": (C.c_int, [C.c_void_p, _fp]),
    "mpm_positions4_wait": (C.c_int, [C.c_void_p]),
    "mpm_density_compute": (C.c_int, [C.c_void_p]),
    "mpm_read_density": (C.c_int, [C.c_void_p, _fp]),
    "mpm_read_density_blocks": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), _fp, C.c_int64,
                                          C.POINTER(C.c_int64)]),
    "mpm_write_density": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mpm_read_sort": (C.c_int, [C.c_void_p, _up, _up, _up, C.c_int64]),
    "mpm_host_keys": (C.c_int, [C.c_void_p, _fp, C.c_int64, _up]),
    "mpm_get_stats": (C.c_int, [C.c_void_p, C.POINTER(MpmStats)]),
    "mpm_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_profile_reset": (C.c_int, [C.c_void_p]),
    "mpm_profile_read": (C.c_int, [C.c_void_p, C.POINTER(MpmKernelTime), C.c_int, C.POINTER(C.c_int)]),
    "mpm_debug_svd3": (C.c_int, [_fp, C.c_int64, _fp, _fp, _fp]),
    "mpm_batch_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "mpm_batch_destroy": (None, [C.c_void_p]),
    "mpm_batch_add": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mpm_batch_remove": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mpm_batch_info": (C.c_int, [C.c_void_p, C.POINTER(MpmBatchStats)]),
    "mpm_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_batch_step_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_batch_synchronize": (C.c_int, [C.c_void_p]),
    "mpm_batch_set_reorder_interval": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_batch_set_incremental_sort": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_batch_read_sort": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "mpm_batch_set_tune": (C.c_int, [C.c_void_p, C.POINTER(MpmTune)]),
    "mpm_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(MpmTune)]),
    "mpm_tune_lookup": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, C.POINTER(MpmTune)]),
    "mpm_batch_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_batch_read_times": (C.c_int, [C.c_void_p, _fp, C.c_int, C.POINTER(C.c_int)]),
    "mpm_batch_health": (C.c_int, [C.c_void_p, C.POINTER(MpmSceneHealth), C.c_int]),
}

_lib = None


def lib_path() -> str:
    return _LIB_PATH


def load(path: str | None = None):
    """Load libmpm_hip.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or _LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"{p} is missing: build it with `python -m mpm_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class MpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mpm error {code}: {msg}")
        self.code = code


def check(rc: int):
    if rc != MPM_OK:
        msg = load().mpm_last_error()
        raise MpmError(rc, msg.decode() if msg else "")


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def uptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_up)
