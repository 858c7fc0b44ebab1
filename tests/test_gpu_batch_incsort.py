"""The batch's cell sort updated from the last step's order: after every
step the batch's sorted order must be exactly the stable sort of the
members' cell keys (scene key offsets included) over the global slots --
through updates, the full sorts it falls back to, graph replays, the fused
transfer and the store reorder of a chunk of members per step; and a batch
stepped with the update must evolve as the same batch sorted in full every
step, up to float-atomic order."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.batch import Batch, Tune
from mpm_amd.solver import Solver
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _members(count, seed, transfer, n_hi=40000):
    rng = scenes.Rng(seed)
    scs = [scenes.batch_scene(rng, i, n_lo=3000, n_hi=n_hi, transfer=transfer) for i in range(count)]
    return scs, [Solver(s.desc, s.particles) for s in scs]


def _check_order(keys, slots, tag):
    """(key, slot) strictly ascending and the slots a permutation."""
    n = keys.size
    k = keys.astype(np.int64)
    s = slots.astype(np.int64)
    asc = np.all((k[1:] > k[:-1]) | ((k[1:] == k[:-1]) & (s[1:] > s[:-1])))
    perm = np.array_equal(np.sort(s), np.arange(n))
    assert asc and perm, (tag, bool(asc), bool(perm))


def _check_keys(keys, slots, descs, positions, tag):
    """Each sorted key is its slot's cell key plus one offset per member,
    the offsets ascending in membership order (positions in store order)."""
    counts = np.array([p.shape[0] for p in positions], np.int64)
    pbase = np.concatenate(([0], np.cumsum(counts)))
    local = [pyoracle.keys(d, p).astype(np.int64) for d, p in zip(descs, positions)]
    m = np.searchsorted(pbase, slots.astype(np.int64), side="right") - 1
    off = np.empty(keys.size, np.int64)
    for j in range(len(positions)):
        sel = m == j
        lk = local[j][slots[sel].astype(np.int64) - pbase[j]]
        d = keys[sel].astype(np.int64) - lk
        assert d.size == counts[j] and np.all(d == d[0]), (tag, j)
        off[sel] = d[0]
    first = [int(off[m == j][0]) for j in range(len(positions))]
    assert all(a < b for a, b in zip(first, first[1:])), (tag, first)


@pytest.mark.parametrize("transfer", [N.APIC, N.PIC_FLIP])
def test_batch_update_is_the_stable_sort(transfer):
    scs, solvers = _members(12, 21, transfer)
    b = Batch(N.QUADRATIC_BSPLINE, transfer)
    for s in solvers:
        b.add(s)
    b.set_reorder_interval(0)
    descs = [s.desc for s in scs]
    # each member's store order (its particles were put in cell order at
    # creation; with the reorder off it stays): the input sequence of a sort
    seqs = []
    for s in solvers:
        s.sort()
        seqs.append(s.read_sort()[2])
    for step in range(8):
        positions = [s.read_particles().position[q] for s, q in zip(solvers, seqs)]
        b.step(1)
        keys, slots = b.read_sort()
        _check_order(keys, slots, step)
        _check_keys(keys, slots, descs, positions, step)
        moved = sum(int((pyoracle.keys(d, s.read_particles().position[q]) != pyoracle.keys(d, p)).sum())
                    for d, s, q, p in zip(descs, solvers, seqs, positions))
        info = b.info()
        print(f"[batch sort {transfer}] step {step + 1}: stable order ok, particles that change cell next={moved} "
              f"of {info['particles']}, incremental sorts so far={info['incremental_sorts']}", flush=True)
    # the first step sorts in full, every later one updates
    assert info["incremental_sorts"] == 7
    b.close()
    for s in solvers:
        s.close()


def test_batch_update_with_fusion_reorder_and_graphs_matches_full_sort():
    runs = []
    for incremental in (True, False):
        scs, solvers = _members(10, 22, N.APIC)
        b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
        for s in solvers:
            b.add(s)
        b.set_tune(Tune(256, 128, 0, fuse_transfer=1))
        b.set_reorder_interval(3)  # a chunk of members renumbered every step
        b.set_incremental_sort(incremental)
        for step in range(6):
            b.step(1)
            keys, slots = b.read_sort()
            _check_order(keys, slots, (incremental, step))
        b.step_graph(6)
        keys, slots = b.read_sort()
        _check_order(keys, slots, (incremental, "graph"))
        info = b.info()
        print(f"[batch sort fused, incremental={incremental}] incremental sorts {info['incremental_sorts']} "
              f"of {info['steps']} steps", flush=True)
        if incremental:
            assert info["incremental_sorts"] >= 10
        else:
            assert info["incremental_sorts"] == 0
        runs.append((b, solvers))
    (ba, sa), (bb, sb) = runs
    worst = 0.0
    for x, y in zip(sa, sb):
        pa, pb = x.read_particles(), y.read_particles()
        assert np.array_equal(pa.mass, pb.mass)
        worst = max(worst, float(np.abs(pa.position.astype(np.float64) - pb.position).max() / x.desc.stride))
    print(f"[batch sort fused] update vs full sort: position difference {worst:.2e} dx after 12 steps", flush=True)
    assert worst <= 1e-4
    for b, solvers in runs:
        b.close()
        for s in solvers:
            s.close()


def test_batch_update_overflow_falls_back():
    # a fast member: ~20% of its 300k particles change cell every step, past
    # the list's capacity (an eighth of the batch's particles): the step
    # sorts in full
    scs, solvers = _members(4, 23, N.APIC)
    fast = scenes.snow_cube(seed=8, n=300_000, res=64, lo=0.3, hi=0.7, velocity=(30.0, 0.0, 0.0), dt=1e-4,
                              transfer=N.APIC, interp=N.QUADRATIC_BSPLINE)
    f = Solver(fast.desc, fast.particles)
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for s in solvers + [f]:
        b.add(s)
    b.set_reorder_interval(0)
    for step in range(3):
        b.step(1)
        keys, slots = b.read_sort()
        _check_order(keys, slots, step)
    assert b.info()["incremental_sorts"] == 0
    state = f.read_particles()
    state.velocity[:] = state.velocity * 0.01
    f.write_particles(state)  # slowed: keys recomputed, one full sort, then updates
    for step in range(3):
        b.step(1)
        keys, slots = b.read_sort()
        _check_order(keys, slots, ("slowed", step))
    assert b.info()["incremental_sorts"] == 2
    b.close()
    for s in solvers + [f]:
        s.close()
