"""Batched scenes (mpm_batch_*): every member of a batch must evolve as the
same scene stepped alone, within float-atomic summation order (1e-5
relative per step), and one member must match the oracle after one step.
"""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.batch import Batch, Tune, tune_lookup
from mpm_amd.solver import Solver, SolverDesc
from oracle import pyoracle
from tests.helpers import particle_errors, report

pytestmark = pytest.mark.gpu

TOL1 = 1e-5


def _scenes(count, seed, n_lo=2000, n_hi=12000, **kw):
    rng = scenes.Rng(seed)
    return [scenes.batch_scene(rng, i, n_lo=n_lo, n_hi=n_hi, **kw) for i in range(count)]


def _errs(pa, pb, h, apic):
    e = particle_errors(pa, pb, ("position", "velocity", "Fe", "Fp"))
    if apic:
        # B ~ v h: relative to max|v| h (B is ~0 for a near-uniform field)
        e["B"] = float(np.abs(pa.B.astype(np.float64) - pb.B).max() / (np.abs(pb.velocity).max() * h + 1e-30))
    return e


def _compare(batch_members, alone, tag, tol, apic=True):
    for k, (m, a) in enumerate(zip(batch_members, alone)):
        e = _errs(m.read_particles(), a.read_particles(), m.desc.stride, apic)
        report(f"{tag} scene {k} (N={m.count})", e)
        assert all(v <= tol for v in e.values()), (k, e)


def test_batch_matches_alone_apic_quadratic():
    scs = _scenes(6, seed=31)
    print("batch scenes:", [(s.particles.count, s.desc.resolution[0]) for s in scs], flush=True)
    alone = [Solver(s.desc, s.particles) for s in scs]
    members = [Solver(s.desc, s.particles) for s in scs]
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for m in members:
        b.add(m)
    info = b.info()
    assert info["scenes"] == 6 and info["particles"] == sum(s.particles.count for s in scs)
    for step in range(3):
        b.step(1)
        for a in alone:
            a.step(1)
        b.synchronize()
        _compare(members, alone, f"batch step {step + 1}", TOL1)
    for m in members:
        assert m.stats()["steps"] == 3


def test_batch_member_matches_oracle_one_step():
    scs = _scenes(3, seed=32, n_lo=1500, n_hi=4000)
    members = [Solver(s.desc, s.particles) for s in scs]
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for m in members:
        b.add(m)
    b.step(1)
    b.synchronize()
    for k, (s, m) in enumerate(zip(scs, members)):
        o = pyoracle.OracleSolver(s.desc, s.particles)
        o.step(1)
        e = _errs(m.read_particles(), o.read_particles(), s.desc.stride, True)
        report(f"batch vs oracle scene {k}", e)
        assert all(v <= TOL1 for v in e.values()), (k, e)


def test_batch_graph_replay_matches_eager():
    scs = _scenes(4, seed=33)
    ea = [Solver(s.desc, s.particles) for s in scs]
    ga = [Solver(s.desc, s.particles) for s in scs]
    be, bg = Batch(N.QUADRATIC_BSPLINE, N.APIC), Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for x, y in zip(ea, ga):
        be.add(x)
        bg.add(y)
    for k in range(3):
        be.step(1)
        bg.step_graph(1)
        bg.synchronize()
        _compare(ga, ea, f"batch graph replay {k + 1}", TOL1)


def test_batch_pic_flip_cubic_escape_and_particle_collision():
    # cubic PIC-FLIP scenes; one has particles outside its grid (escape
    # keys), one has particle collision switched on
    scs = _scenes(3, seed=34, transfer=N.PIC_FLIP, interp=N.CUBIC_BSPLINE)
    st = scs[0].particles
    pos = st.position.copy()
    pos[:50] += 2.0  # far outside the unit domain
    scs[0].particles.position[:] = pos
    alone = [Solver(s.desc, s.particles) for s in scs]
    members = [Solver(s.desc, s.particles) for s in scs]
    alone[1].set_particles_collision(True)
    members[1].set_particles_collision(True)
    b = Batch(N.CUBIC_BSPLINE, N.PIC_FLIP)
    for m in members:
        b.add(m)
    for step in range(2):
        b.step(1)
        for a in alone:
            a.step(1)
        b.synchronize()
        _compare(members, alone, f"cubic PIC-FLIP step {step + 1}", TOL1, apic=False)
    assert members[0].stats()["escaped"] == alone[0].stats()["escaped"] > 0


def test_batch_membership_changes():
    scs = _scenes(5, seed=35)
    alone = [Solver(s.desc, s.particles) for s in scs]
    members = [Solver(s.desc, s.particles) for s in scs]
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for m in members[:3]:
        b.add(m)
    b.step(2)
    for a in alone[:3]:
        a.step(2)
    # scene 1 leaves and keeps stepping alone; scenes 3, 4 join
    b.remove(members[1])
    members[1].step(1)
    alone[1].step(1)
    b.add(members[3])
    b.add(members[4])
    b.step(2)
    for k in (0, 2, 3, 4):
        alone[k].step(2)
    # a member stepped on its own while in the batch, then with the batch
    members[0].step(1)
    alone[0].step(1)
    b.step(1)
    for k in (0, 2, 3, 4):
        alone[k].step(1)
    b.synchronize()
    _compare(members, alone, "membership", TOL1)
    # destroying a member removes it from the batch
    members[2].close()
    assert b.info()["scenes"] == 3
    b.step(1)
    b.synchronize()


def test_batch_rejects_mismatched_scene():
    s = _scenes(1, seed=36)[0]
    b = Batch(N.QUADRATIC_BSPLINE, N.PIC_FLIP)
    m = Solver(s.desc, s.particles)  # APIC scene
    with pytest.raises(N.MpmError):
        b.add(m)
    b2 = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    b2.add(m)
    with pytest.raises(N.MpmError):
        b2.add(m)
    with pytest.raises(N.MpmError):
        b.set_tune(Tune(item_max=100))


def test_batch_autotune_and_fixed_tune():
    scs = _scenes(4, seed=37)
    alone = [Solver(s.desc, s.particles) for s in scs]
    members = [Solver(s.desc, s.particles) for s in scs]
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for m in members:
        b.add(m)
    t = b.autotune(trial_steps=1)
    print("autotune choice:", t, flush=True)
    assert t.item_max in (128, 256, 512, 1024) and t.g2p_threads in (64, 128, 256) and t.step_ms > 0
    total = sum(s.particles.count for s in scs)
    assert tune_lookup(N.QUADRATIC_BSPLINE, N.APIC, total) is not None
    b.set_tune(Tune(item_max=128, g2p_threads=256, workgroups=300))
    b.step(1)
    b.synchronize()
    steps = members[0].stats()["steps"]
    assert all(m.stats()["steps"] == steps for m in members)
    for a in alone:
        a.step(steps)
    # many steps of differently ordered float atomics: 1e-4 relative
    _compare(members, alone, f"after autotune ({steps} steps)", 1e-4)


def test_batch_timing_records_steps():
    scs = _scenes(2, seed=38)
    members = [Solver(s.desc, s.particles) for s in scs]
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for m in members:
        b.add(m)
    b.timing(True)
    b.step(3)
    b.step_graph(2)
    t = b.read_times()
    assert t.shape == (5,) and (t > 0).all()
    assert b.read_times().shape == (0,)
