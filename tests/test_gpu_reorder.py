"""Store reorder (mpm_set_reorder_interval, mpm_batch_set_reorder_interval):
the particle store is permuted in place into cell order every few steps.
Results must equal a run that never reorders, within float-atomic order
(1e-5 relative), in eager, graph and batched stepping; reads stay in the
caller's order, and the sort stays bit-exact against std::stable_sort."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.batch import Batch
from mpm_amd.solver import Solver
from oracle import pyoracle
from tests.helpers import particle_errors, particle_errors_apic, rel_err, report

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _pair(sc, interval):
    a = Solver(sc.desc, sc.particles)
    b = Solver(sc.desc, sc.particles)
    a.set_reorder_interval(interval)
    b.set_reorder_interval(0)
    return a, b


@pytest.mark.parametrize("transfer", [N.APIC, N.PIC_FLIP])
def test_reorder_eager_matches_no_reorder(transfer):
    sc = scenes.snow_cube(seed=41, n=20000, res=48, transfer=transfer, velocity=(2.0, -3.0, 1.0), dt=2e-4)
    a, b = _pair(sc, 3)
    print("reorder eager", transfer, flush=True)
    a.step(20)
    b.step(20)
    pa, pb = a.read_particles(), b.read_particles()
    e = particle_errors_apic(pa, pb, sc.desc.stride) if transfer == N.APIC else particle_errors(pa, pb)
    e["positions4"] = rel_err(a.read_positions4(), b.read_positions4())
    e["density"] = rel_err(a.density(), b.density())
    report(f"reorder every 3 steps ({'APIC' if transfer else 'PIC-FLIP'})", e)
    assert all(v <= TOL for v in e.values()), e
    assert a.stats()["steps"] == 20


def test_reorder_between_graph_replays():
    sc = scenes.snow_cube(seed=42, n=15000, res=48, transfer=N.APIC, velocity=(-1.0, 2.0, 0.5), dt=2e-4)
    a, b = _pair(sc, 4)
    a.step_graph(13)
    b.step(13)
    e = particle_errors_apic(a.read_particles(), b.read_particles(), sc.desc.stride)
    report("reorder between graph replays", e)
    assert all(v <= TOL for v in e.values()), e
    assert a.stats()["steps"] == 13


def test_sort_bit_exact_after_reorders():
    sc = scenes.snow_cube(seed=43, n=12000, res=32, transfer=N.APIC, velocity=(3.0, -2.0, 1.0), dt=5e-4)
    gpu = Solver(sc.desc, sc.particles)
    gpu.set_reorder_interval(1)
    gpu.step(6)
    gpu.sort()
    keys, ids, seq = gpu.read_sort()
    pos = gpu.read_particles().position
    k_seq = pyoracle.keys(sc.desc, pos[seq])
    order = pyoracle.stable_argsort(k_seq)
    assert np.array_equal(keys, k_seq[order])
    assert np.array_equal(ids, seq[order])
    # the store is in cell order: the sort input is sorted but for the
    # particles that changed cell in the last step (here ~10% per step:
    # |v| dt / h summed over the axes is 0.1), against ~50% descents for a
    # store in random order
    assert np.count_nonzero(np.diff(k_seq.astype(np.int64)) < 0) < 0.15 * keys.size


def test_batch_reorder_matches_alone():
    rng = scenes.Rng(44)
    scs = [scenes.batch_scene(rng, i, n_lo=2000, n_hi=9000) for i in range(5)]
    alone = [Solver(x.desc, x.particles) for x in scs]
    members = [Solver(x.desc, x.particles) for x in scs]
    for s in alone:
        s.set_reorder_interval(0)
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    b.set_reorder_interval(2)
    for m in members:
        b.add(m)
    b.step(5)
    b.step_graph(4)
    for s in alone:
        s.step(9)
    b.synchronize()
    worst = 0.0
    for m, s in zip(members, alone):
        e = particle_errors_apic(m.read_particles(), s.read_particles(), m.desc.stride)
        worst = max(worst, max(e.values()))
        assert all(v <= TOL for v in e.values()), e
    print(f"[batch reorder every 2 steps] worst {worst:.3e}", flush=True)
    # a member leaves and keeps stepping on its own store order
    b.remove(members[0])
    members[0].step(2)
    alone[0].step(2)
    e = particle_errors(members[0].read_particles(), alone[0].read_particles())
    report("member after leaving a reordering batch", e)
    assert all(v <= TOL for v in e.values()), e
