"""Render export (replacement of updateGLBufferByCPU / updateGLBufferWithCUDA,
src/mpm_solver3d.cu:457-508): float4 (x, y, z, 1) positions in the original
particle order, packed into a caller's device buffer, or copied to host
memory on a second stream while later steps run.  Integer-exact: every
export must hold the same bits as the synchronous readback of that frame.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the solver library links (same process-wide instance)."""
    hip = C.CDLL("libamdhip64.so.7")
    for f in ("hipHostMalloc", "hipMalloc"):
        getattr(hip, f).restype = C.c_int
    hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    hip.hipHostFree.argtypes = [C.c_void_p]
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _solver(seed=31, n=20000):
    sc = scenes.snow_cube(seed=seed, n=n, res=64, transfer=N.APIC, velocity=(0.5, -0.5, 0.2))
    return sc, Solver(sc.desc, sc.particles)


def test_async_export_matches_readback_and_overlaps_steps():
    sc, s = _solver()
    s.step(2)
    frame = s.read_positions4()
    assert np.array_equal(frame[:, :3], s.read_particles().position) and np.all(frame[:, 3] == 1.0)
    out = np.full((s.count, 4), np.nan, np.float32)
    s.positions4_async(out)
    s.step(5)  # queued behind the pack; the copy runs on the export stream
    s.positions4_wait()
    assert np.array_equal(out, frame)
    # a second export (after more steps) reuses the staging buffer
    out2 = np.zeros_like(out)
    s.positions4_async(out2)
    s.positions4_wait()
    assert np.array_equal(out2, s.read_positions4())
    assert not np.array_equal(out2, frame)


def test_async_export_into_pinned_memory():
    hip = _hip()
    sc, s = _solver(seed=32, n=50000)
    s.step(3)
    nbytes = s.count * 16
    ptr = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(ptr), nbytes, 0) == 0
    try:
        view = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(s.count, 4))
        view[:] = np.nan
        s.positions4_async(view)
        s.step(2)
        s.positions4_wait()
        s2 = Solver(sc.desc, sc.particles)
        s2.step(3)
        ref = s2.read_positions4()
        # an identical twin stepped 3 times: same frame within float-atomic order
        assert np.abs(view - ref).max() <= 1e-5 * np.abs(ref).max()
        assert np.all(view[:, 3] == 1.0)
    finally:
        hip.hipHostFree(ptr)


def test_pack_into_device_buffer():
    hip = _hip()
    sc, s = _solver(seed=33, n=10000)
    s.step(1)
    nbytes = s.count * 16
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), nbytes) == 0
    try:
        s.pack_positions4(dev.value)
        s.synchronize()
        host = np.full((s.count, 4), np.nan, np.float32)
        assert hip.hipMemcpy(host.ctypes.data, dev, nbytes, 2) == 0  # hipMemcpyDeviceToHost
        assert np.array_equal(host, s.read_positions4())
    finally:
        hip.hipFree(dev)
