"""100-step accuracy on every workload shape: the device state after 100
steps against the FP32 oracle (mass exact, max position difference <=
1e-3 * dx), with the distance to the FP64 oracle reported beside it.  The
shapes are the benchmark workloads at sizes the serial oracle steps in
seconds: two colliding APIC balls (cubic), the APIC slab with random
velocities (quadratic), members of a batch (store reorders every 16 steps
included) and requests of the served stream (PIC-FLIP and APIC)."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.batch import Batch
from mpm_amd.serve import Server, estimate_bytes, requests_from_scenes
from mpm_amd.solver import Solver
from oracle import pyoracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

STEPS = 100


def _check(tag, desc, particles, got):
    """got: the device state after STEPS steps (caller's order)."""
    r32 = pyoracle.OracleSolver(desc, particles, 32)
    r64 = pyoracle.OracleSolver(desc, particles, 64)
    r32.step(STEPS)
    r64.step(STEPS)
    r = r32.read_particles()
    q = r64.read_particles64()
    dx = desc.stride
    drift = float(np.abs(got.position.astype(np.float64) - r.position).max())
    drift64 = float(np.abs(got.position.astype(np.float64) - q["position"]).max())
    o32_64 = float(np.abs(r.position.astype(np.float64) - q["position"]).max())
    print(f"[{tag} x{STEPS}] N={particles.count} drift vs FP32 oracle {drift / dx:.2e} dx, vs FP64 oracle "
          f"{drift64 / dx:.2e} dx (FP32 oracle vs FP64 {o32_64 / dx:.2e} dx); velocity rel "
          f"{rel_err(got.velocity, r.velocity):.2e}", flush=True)
    assert np.array_equal(got.mass, r.mass)
    assert float(got.mass.astype(np.float64).sum()) == float(particles.mass.astype(np.float64).sum())
    assert drift <= 1e-3 * dx, (tag, drift / dx)


def test_two_balls_apic_cubic():
    sc = scenes.two_balls(seed=11, res=32, ppc=4)
    s = Solver(sc.desc, sc.particles)
    s.step(STEPS)
    _check("two_balls_apic_cubic", sc.desc, sc.particles, s.read_particles())


def test_slab_apic_quadratic():
    sc = scenes.apic_slab(seed=12, nx=24, ny=12, nz=24, res=32)
    s = Solver(sc.desc, sc.particles)
    s.step(STEPS)
    _check("slab_apic_quadratic", sc.desc, sc.particles, s.read_particles())


def test_batch_members():
    rng = scenes.Rng(13)
    scs = [scenes.batch_scene(rng, i, n_lo=2000, n_hi=6000) for i in range(3)]
    members = [Solver(x.desc, x.particles) for x in scs]
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for m in members:
        b.add(m)
    b.step(STEPS)
    b.synchronize()
    for i, (x, m) in enumerate(zip(scs, members)):
        _check(f"batch member {i}", x.desc, x.particles, m.read_particles())


def test_served_requests():
    stream = scenes.request_stream(seed=14, rate=50.0, duration=0.1, n_lo=2000, n_hi=6000, steps_lo=STEPS,
                                   steps_hi=STEPS, max_requests=4)
    reqs = requests_from_scenes(stream)
    assert reqs and {r.desc.transfer for r in reqs} == {N.PIC_FLIP, N.APIC}, [r.desc.transfer for r in reqs]
    srv = Server(budget_bytes=8 * max(estimate_bytes(r.desc, r.particles.count) for r in reqs), tick_steps=16,
                 keep_results=True)
    try:
        rep = srv.run(reqs)
    finally:
        srv.close()
    recs = {r.rid: r for r in rep.records}
    for r in reqs:
        kind = "APIC" if r.desc.transfer == N.APIC else "PIC-FLIP"
        _check(f"served rid={r.rid} {kind}", r.desc, r.particles, recs[r.rid].result)
