"""Edge cases of the device path: one particle (closed-form free fall),
an empty scene (alone and inside a batch), and particles outside the grid
against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.batch import Batch
from mpm_amd.solver import ParticleInput, Solver, SolverDesc
from oracle import pyoracle
from tests.helpers import particle_errors, report

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_single_particle_free_fall(transfer):
    # one interior particle, no neighbours: grad v = 0 so F stays I and the
    # velocity is -g n dt, y_n = y0 - g dt^2 n (n+1) / 2 (symplectic Euler);
    # the reference's mass threshold keeps fringe nodes at v* = 0, which is
    # exact here because m = 1 kg makes every stencil node heavy
    d = SolverDesc(resolution=(32, 32, 32), stride=1 / 32, transfer=transfer, dt=1e-3, gravity=9.8,
                   interp=N.QUADRATIC_BSPLINE)
    p = ParticleInput(np.array([[0.5, 0.6, 0.4]]), np.zeros((1, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    s = Solver(d, p)
    n = 20
    s.step(n)
    st = s.read_particles()
    print("free fall:", st.position[0], st.velocity[0], flush=True)
    np.testing.assert_allclose(st.velocity[0], [0, -9.8 * n * 1e-3, 0], atol=2e-5)
    np.testing.assert_allclose(st.position[0, 1], 0.6 - 9.8 * 1e-6 * n * (n + 1) / 2, atol=2e-6)
    np.testing.assert_allclose(st.Fe[0], np.eye(3), atol=2e-6)
    assert s.stats()["steps"] == n


def test_empty_scene_alone_and_in_batch():
    d = SolverDesc(resolution=(16, 16, 16), stride=1 / 16, transfer=N.APIC)
    empty = ParticleInput(np.zeros((0, 3)), np.zeros((0, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    s = Solver(d, empty)
    s.step(3)
    s.step_graph(2)
    assert s.read_particles().count == 0
    assert s.stats()["steps"] == 5 and s.stats()["active_nodes"] == 0
    # an empty member beside a real one: the real one is unaffected
    sc = scenes.snow_cube(seed=12, n=3000, res=32, transfer=N.APIC, velocity=(0.2, 0.0, -0.1))
    alone = Solver(sc.desc, sc.particles)
    member = Solver(sc.desc, sc.particles)
    e2 = Solver(d, empty)
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    b.add(e2)
    b.add(member)
    b.step(3)
    alone.step(3)
    b.synchronize()
    e = particle_errors(member.read_particles(), alone.read_particles())
    report("batch with an empty scene", e)
    assert all(v <= 1e-5 for v in e.values()), e
    assert e2.stats()["steps"] == 3


@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_particles_outside_grid_match_oracle(transfer):
    # a quarter of the particles start outside the grid (stencil misses it:
    # the escape key); inside ones must match the oracle, outside ones follow
    # the reference (no grid contribution: PIC-FLIP keeps the FLIP part of v,
    # APIC gets v = 0; F unchanged apart from the clamp)
    sc = scenes.snow_cube(seed=13, n=4000, res=32, transfer=transfer, velocity=(0.5, -0.3, 0.2))
    pos = sc.particles.position.copy()
    pos[::4] += np.float32(1.5)
    sc.particles.position[:] = pos
    gpu = Solver(sc.desc, sc.particles)
    ref = pyoracle.OracleSolver(sc.desc, sc.particles, 32)
    assert gpu.stats()["escaped"] == 1000
    gpu.step(2)
    ref.step(2)
    e = particle_errors(gpu.read_particles(), ref.read_particles())
    report(f"escaped quarter ({'APIC' if transfer else 'PIC-FLIP'})", e)
    assert all(v <= 1e-5 for v in e.values()), e


def test_cfl_monitor():
    # free fall: every particle has |v| = g n dt exactly (see above), so
    # max_speed and cfl = max_speed dt / h have closed forms
    d = SolverDesc(resolution=(32, 32, 32), stride=1 / 32, transfer=N.APIC, dt=1e-3, gravity=9.8,
                   interp=N.QUADRATIC_BSPLINE)
    p = ParticleInput(np.array([[0.5, 0.6, 0.4], [0.3, 0.7, 0.6]]), np.zeros((2, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025,
                      0.0075, 0)
    s = Solver(d, p)
    st0 = s.stats()
    assert st0["max_speed"] == 0.0 and st0["cfl"] == 0.0
    s.step(10)
    st = s.stats()
    v = 9.8 * 10 * 1e-3
    assert st["max_speed"] == pytest.approx(v, rel=2e-5)
    assert st["cfl"] == pytest.approx(v * 1e-3 * 32, rel=2e-5)
    # a NaN velocity is reported, not hidden
    state = s.read_particles()
    state.velocity[1, 0] = np.nan
    s.write_particles(state)
    assert np.isnan(s.stats()["max_speed"])


@pytest.mark.parametrize("fused", [False, True], ids=["two-pass", "fused"])
def test_profile_levels(fused):
    # level 1 times every launch group, level 2 only the item kernels and the
    # store reorder; other levels are rejected.  With transfer fusion the
    # first step after creation runs P2G alone, every step the fused kernel.
    sc = scenes.snow_cube(seed=9, n=5000, res=32, transfer=N.APIC)
    s = Solver(sc.desc, sc.particles)
    s.set_transfer_fusion(fused)
    item = "g2p2g" if fused else "g2p"
    s.profile(1)
    s.profile_reset()
    s.step(3)
    full = s.profile_read()
    assert {"p2g", item, "sort_hist", "sort_scatter", "items", "grid_clear"} <= set(full), full
    assert full[item][0] == 3 and full["p2g"][0] == (1 if fused else 3)
    s.profile(2)
    s.profile_reset()
    s.set_reorder_interval(2)
    s.step(4)
    items = s.profile_read()
    assert set(items) == ({"g2p2g", "reorder"} if fused else {"p2g", "g2p", "reorder"}), items
    assert items[item][0] == 4 and items[item][1] > 0.0
    with pytest.raises(Exception):
        s.profile(3)
    s.profile(False)
    s.profile_reset()
    s.step(1)
    assert s.profile_read() == {}
    # switching levels without a reset: the skipped kernels' end events must
    # not re-close the pairs the previous level opened
    s.profile(1)
    s.step(1)
    s.profile(2)
    s.step(3)
    mixed = s.profile_read()
    s.profile(False)
    one = full["sort_hist"][1] / full["sort_hist"][0]
    assert mixed[item][0] == 4 and mixed["sort_hist"][0] >= 1  # the level-1 step's passes only
    assert mixed["sort_hist"][1] / mixed["sort_hist"][0] < 20 * one + 0.05, (mixed["sort_hist"], one)
