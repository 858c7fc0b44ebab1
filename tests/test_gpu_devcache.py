"""Device block cache: a scene built from reused blocks (which still hold
an earlier scene's data) steps like the oracle, and building or destroying
a scene does not wait for the work other streams have queued.

Tolerance (FP32): 1e-5 normwise relative against the oracle after 2
steps, as for every one-step parity check."""
import time

import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.solver import Solver, cache_info, release_cache, set_cache_limit
from oracle import pyoracle
from tests.helpers import particle_errors_apic, report

pytestmark = pytest.mark.gpu

TOL = 1e-5


def test_reused_blocks_match_oracle():
    # same particle count and grid, other content: its blocks fit the next scene
    junk = scenes.snow_cube(seed=3, n=16384, res=64, lo=0.2, hi=0.8, transfer=N.PIC_FLIP,
                            interp=N.QUADRATIC_BSPLINE, velocity=(-5.0, 3.0, 1.0), dt=1e-4)
    s = Solver(junk.desc, junk.particles)
    s.step(5)
    s.synchronize()
    s.close()
    before = cache_info()
    assert before["cached_bytes"] > 0 and before["live_bytes"] == 0, before
    sc = scenes.snow_cube(seed=11, n=16384, res=64, lo=0.35, hi=0.65, transfer=N.APIC,
                          interp=N.QUADRATIC_BSPLINE, velocity=(10.0, -4.0, 2.0), dt=1e-4)
    gpu = Solver(sc.desc, sc.particles)
    after = cache_info()
    print(f"[cache] before {before}, after build {after}", flush=True)
    assert after["hits"] - before["hits"] >= 10  # the grid, sort buffers and particle store
    ref = pyoracle.OracleSolver(sc.desc, sc.particles, 32)
    for k in range(2):
        gpu.step(1)
        ref.step(1)
        e = particle_errors_apic(gpu.read_particles(), ref.read_particles(), sc.desc.stride)
        report(f"reused blocks, step {k + 1}", e)
        assert all(v <= TOL for v in e.values()), (k, e)
    gpu.close()
    ref.close()


def test_build_and_destroy_do_not_wait_for_other_streams():
    slab = scenes.apic_slab(seed=2)  # 8.4M particles: 100 steps are ~0.15 s of queued work
    big = Solver(slab.desc, slab.particles)
    big.step(2)
    big.synchronize()
    small = scenes.snow_cube(seed=5, n=20000, res=64)
    Solver(small.desc, small.particles).close()  # its sizes are cached now
    t0 = time.perf_counter()
    big.step(100)
    t1 = time.perf_counter()
    s = Solver(small.desc, small.particles)
    s.close()
    t2 = time.perf_counter()
    big.synchronize()
    t3 = time.perf_counter()
    print(f"[overlap] enqueue {1e3 * (t1 - t0):.1f} ms, build + destroy {1e3 * (t2 - t1):.1f} ms, "
          f"queued work left afterwards {1e3 * (t3 - t2):.1f} ms", flush=True)
    # hipFree would have drained the slab's queue before returning
    assert t3 - t2 > 0.02 and t2 - t1 < t3 - t1, (t1 - t0, t2 - t1, t3 - t2)
    big.close()


def test_cache_limit_and_release():
    sc = scenes.snow_cube(seed=6, n=8192, res=32)
    Solver(sc.desc, sc.particles).close()
    assert cache_info()["cached_bytes"] > 0
    try:
        set_cache_limit(0)  # trims at once; later releases go straight back
        assert cache_info()["cached_bytes"] == 0
        Solver(sc.desc, sc.particles).close()
        assert cache_info()["cached_bytes"] == 0
    finally:
        set_cache_limit(16 << 30)
    Solver(sc.desc, sc.particles).close()
    assert cache_info()["cached_bytes"] > 0
    release_cache()
    info = cache_info()
    assert info["cached_bytes"] == 0, info
    with pytest.raises(N.MpmError):
        set_cache_limit(-1)
