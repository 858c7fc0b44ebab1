"""The benchmark's 8.4M-particle APIC slab (256^3 grid) at full size, through
properties that do not need the serial oracle to step it: the sort of all
8.4M cell keys bit-exact against a stable argsort of independently computed
keys; P2G conservation of mass and momentum (every stencil is inside the
grid: sum_i m_i = sum_p m_p, and sum_i p_i = sum_p m_p v_p since the APIC
term sums to zero); and after 50 steps (with the store reorder every 100
steps not yet due, then one forced) particle masses unchanged bit for bit,
finite state, positions inside the domain."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import scenes
from mpm_amd.solver import Solver
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slab():
    return scenes.workload("apic_slab")


def test_full_slab_sort_bit_exact(slab):
    s = Solver(slab.desc, slab.particles)
    s.step(3)
    s.sort()
    keys, ids, seq = s.read_sort()
    pos = s.read_particles().position
    k_seq = pyoracle.keys(slab.desc, pos[seq])
    order = pyoracle.stable_argsort(k_seq)
    print(f"[full slab sort] N={keys.size} distinct keys={np.unique(keys).size}", flush=True)
    assert np.array_equal(keys, k_seq[order])
    assert np.array_equal(ids, seq[order])
    s.close()


def test_full_slab_conservation(slab):
    p = slab.particles
    s = Solver(slab.desc, p)
    s.set_grid_record(True)  # momentum and force in their own planes
    s.step(1)
    g = s.read_grid()
    m = g["mass"].astype(np.float64)
    heavy = m > np.finfo(np.float32).eps
    mom = np.where(heavy[:, None], g["velocity"].astype(np.float64) * m[:, None], g["velocity"])
    m_tot = float(p.mass.astype(np.float64).sum())
    p_tot = (p.mass.astype(np.float64)[:, None] * p.velocity.astype(np.float64)).sum(0)
    p_scale = float((p.mass.astype(np.float64)[:, None] * np.abs(p.velocity.astype(np.float64))).sum())
    e_m = abs(m.sum() - m_tot) / m_tot
    e_p = float(np.abs(mom.sum(0) - p_tot).max()) / p_scale
    print(f"[full slab P2G] mass rel err {e_m:.2e}, momentum err / sum m|v| {e_p:.2e}, "
          f"active nodes {int(heavy.sum())}", flush=True)
    assert e_m <= 1e-5 and e_p <= 1e-5
    s.close()


def test_full_slab_fifty_steps(slab):
    p = slab.particles
    s = Solver(slab.desc, p)
    s.step(50)
    s.set_reorder_interval(1)
    s.step(1)  # one store reorder on the full store
    st = s.read_particles()
    stats = s.stats()
    print(f"[full slab x51] max |v| {stats['max_speed']:.3f} m/s, CFL {stats['cfl']:.3f}, "
          f"active nodes {stats['active_nodes_sum'] / max(stats['steps'], 1):.0f} per step", flush=True)
    assert np.array_equal(st.mass, p.mass)
    for f in ("position", "velocity", "Fe", "Fp", "B"):
        assert np.isfinite(getattr(st, f)).all(), f
    assert (st.position > 0.0).all() and (st.position < 1.0).all()
    assert np.isfinite(stats["cfl"]) and stats["cfl"] < 1.0
    s.close()
