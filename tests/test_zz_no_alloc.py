"""The allocation-free readback (mpm_read_particles_noalloc) against the
allocating one (mpm_read_particles), through the C-ABI: bit-identical
state on two calls in a row, the second on a smaller handle, for APIC and
PIC-FLIP, and a clear error above the array's size."""
import ctypes as C

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.solver import ParticleState, Solver

pytestmark = pytest.mark.gpu

FIELDS = ("volume0", "mass", "position", "velocity", "Fe", "Fp", "mu0", "lambda0", "hardening",
          "critical_compression", "critical_stretch", "model", "B")


def _read(s, fn):
    st = ParticleState.empty(s.count)
    c = st.to_c()
    rc = fn(s._h, C.byref(c))
    N.check(rc)
    return st


def _same(a, b, tag):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert np.array_equal(x, y), (tag, f)
    print(f"[{tag}] {a.count} particles: every field bit-identical", flush=True)


@pytest.mark.parametrize("transfer", [N.APIC, N.PIC_FLIP])
def test_noalloc_read_matches_read_twice(transfer):
    lib = N.load()
    big = scenes.snow_cube(seed=21, n=200_000, res=64, transfer=transfer, velocity=(0.3, -0.5, 0.1))
    small = scenes.snow_cube(seed=22, n=3_000, res=32, transfer=transfer, velocity=(-0.2, 0.4, 0.0))
    for k, sc in enumerate((big, small)):  # the second call on a smaller handle
        s = Solver(sc.desc, sc.particles)
        s.step(3)
        got = _read(s, lib.mpm_read_particles_noalloc)
        want = _read(s, lib.mpm_read_particles)
        _same(got, want, f"{'APIC' if transfer == N.APIC else 'PIC-FLIP'} call {k + 1}")
        assert s.stats()["steps"] == 3
        s.close()


def test_noalloc_read_rejects_more_than_its_array():
    sc = scenes.snow_cube(seed=23, n=(1 << 18) + 1, res=64)
    s = Solver(sc.desc, sc.particles)
    with pytest.raises(N.MpmError) as ei:
        s.read_particles(noalloc=True)
    assert ei.value.code == N.MPM_ERR_UNSUPPORTED and "262144" in str(ei.value)
    # the allocating readback still serves it, and a small handle still reads after the refusal
    assert s.read_particles().count == (1 << 18) + 1
    s.close()
    sc = scenes.snow_cube(seed=24, n=1000, res=32)
    s = Solver(sc.desc, sc.particles)
    _same(s.read_particles(noalloc=True), s.read_particles(), "after a refusal")
    s.close()
