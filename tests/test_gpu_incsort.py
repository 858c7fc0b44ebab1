"""The fused step's cell sort of a large scene (>= 2^20 particles), which
updates the previous step's order when few particles changed cell: after
every step the permutation that step used equals std::stable_sort of the
cell keys of the positions the step started from -- through incremental
steps, a move list that overflows (full sorts), store reorders, a state
write and hipGraph replays."""
import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.solver import Solver
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _expected(desc, gpu, pos):
    """Keys and ids of the stable sort of the step's start positions, in the
    sort's input order (the particle store's slots)."""
    _, _, seq = gpu.read_sort()
    k_seq = pyoracle.keys(desc, pos[seq])
    order = pyoracle.stable_argsort(k_seq)
    return k_seq[order], seq[order]


def _step_and_check(gpu, desc, tag, graph=False):
    pos = gpu.read_particles().position
    if graph:
        gpu.step_graph(1)
    else:
        gpu.step(1)
    keys, ids, _ = gpu.read_sort()
    ref_keys, ref_ids = _expected(desc, gpu, pos)
    ok = np.array_equal(keys, ref_keys) and np.array_equal(ids, ref_ids)
    moved = int((pyoracle.keys(desc, gpu.read_particles().position) != pyoracle.keys(desc, pos)).sum())
    st = gpu.stats()
    print(f"[{tag}] step {st['steps']}: bit-exact={ok} particles that change cell next={moved} "
          f"incremental sorts so far={st['incremental_sorts']}", flush=True)
    assert ok, tag
    return st


def test_incremental_sort_bit_exact_slab():
    # 1,048,576 particles, random velocities: ~0.1% change cell per step
    sc = scenes.apic_slab(seed=7, nx=128, ny=64, nz=128, res=128, dt=5e-5, sigma_v=2.0)
    gpu = Solver(sc.desc, sc.particles)
    gpu.set_reorder_interval(6)
    for k in range(14):
        st = _step_and_check(gpu, sc.desc, "slab eager")
    # first step and the steps after the reorders (6, 12) sort in full
    assert st["incremental_sorts"] >= 9
    # a run of updates: the full sort's launches switch to fewer passes on
    # fewer workgroups, which the full sort after the state write uses
    gpu.set_reorder_interval(0)
    for k in range(24):
        st = _step_and_check(gpu, sc.desc, "slab updates")
    # a state write: keys recomputed, the next sort in full, then updates again
    s0 = st["incremental_sorts"]
    state = gpu.read_particles()
    state.velocity[::3] *= -1.0
    gpu.write_particles(state)
    for k in range(3):
        st = _step_and_check(gpu, sc.desc, "after write")
    assert st["incremental_sorts"] == s0 + 2
    # graph replays (a reorder between replays every 4 steps)
    gpu.set_reorder_interval(4)
    for k in range(7):
        st = _step_and_check(gpu, sc.desc, "slab graph", graph=True)
    assert st["incremental_sorts"] > s0 + 4
    gpu.close()


def test_incremental_sort_overflow_falls_back():
    # a fast translation: ~20% of 1.1M particles change cell every step,
    # past the move list's 16,384 entries, so every step sorts in full
    sc = scenes.snow_cube(seed=8, n=1_100_000, res=64, lo=0.3, hi=0.7, velocity=(30.0, 0.0, 0.0), dt=1e-4)
    gpu = Solver(sc.desc, sc.particles)
    for k in range(3):
        st = _step_and_check(gpu, sc.desc, "overflow")
    assert st["incremental_sorts"] == 0
    # slowed down: updates take over
    state = gpu.read_particles()
    state.velocity[:] = state.velocity * 0.01
    gpu.write_particles(state)
    for k in range(3):
        st = _step_and_check(gpu, sc.desc, "slowed")
    assert st["incremental_sorts"] == 2
    gpu.close()


def test_incremental_path_matches_full_sort_run():
    """The same slab stepped with the update and with the full sort each step
    (a forced reorder every step keeps the update off): equal particle
    state after 10 steps up to float-atomic order."""
    sc = scenes.apic_slab(seed=9, nx=128, ny=64, nz=128, res=128, dt=5e-5, sigma_v=2.0)
    a = Solver(sc.desc, sc.particles)
    b = Solver(sc.desc, sc.particles)
    a.set_reorder_interval(0)
    b.set_reorder_interval(1)
    a.step(10)
    b.step(10)
    sa, sb = a.stats(), b.stats()
    pa, pb = a.read_particles(), b.read_particles()
    dx = sc.desc.stride
    d = float(np.abs(pa.position.astype(np.float64) - pb.position).max() / dx)
    print(f"[update vs full] incremental sorts {sa['incremental_sorts']} vs {sb['incremental_sorts']}, "
          f"position difference {d:.2e} dx", flush=True)
    assert sa["incremental_sorts"] == 9 and sb["incremental_sorts"] == 0
    assert np.array_equal(pa.mass, pb.mass) and d <= 1e-4
    a.close()
    b.close()
