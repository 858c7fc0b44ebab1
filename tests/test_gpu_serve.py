"""Served request stream on the device: requests that arrive while others
run are admitted under the HBM budget and join the running batches, and
each result equals the same request run alone.  The comparison is after
tens of steps, so it uses the many-step tolerance: mass exact, position
difference <= 1e-3 * dx, and Fe within 1e-3 relative."""
from __future__ import annotations

import json

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.serve import Server, estimate_bytes, requests_from_scenes
from mpm_amd.solver import Solver
from tests.helpers import rel_err, report

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    stream = scenes.request_stream(seed=41, rate=40.0, duration=0.3, n_lo=2000, n_hi=9000, steps_lo=20,
                                   steps_hi=60, max_requests=8)
    reqs = requests_from_scenes(stream)
    print("requests:", [(r.rid, r.particles.count, r.steps, r.desc.transfer, round(r.arrival, 3)) for r in reqs],
          flush=True)
    assert len(reqs) >= 4
    # budget for about three resident scenes: later requests queue
    budget = 3 * max(estimate_bytes(r.desc, r.particles.count) for r in reqs)
    srv = Server(budget_bytes=budget, tick_steps=8, keep_results=True)
    try:
        rep = srv.run(reqs)
    finally:
        srv.close()
    print("serve report:", rep.summary(), flush=True)
    assert rep.completed == len(reqs)
    assert rep.peak_bytes <= budget
    recs = {r.rid: r for r in rep.records}
    # continuous batching: some request was admitted while another one ran
    overlap = any(a.admitted < b.admitted < a.finished for a in recs.values() for b in recs.values() if a is not b)
    assert overlap
    for r in reqs:
        alone = Solver(r.desc, r.particles)
        alone.step(r.steps)
        pa = alone.read_particles()
        alone.close()
        pm = recs[r.rid].result
        dx = r.desc.stride
        e = {
            "pos/dx": float(np.abs(pm.position.astype(np.float64) - pa.position).max() / dx),
            "Fe": rel_err(pm.Fe, pa.Fe),
            "mass": float(np.abs(pm.mass.astype(np.float64) - pa.mass).max()),
        }
        report(f"served vs alone rid={r.rid} N={r.particles.count} steps={r.steps}", e)
        assert e["mass"] == 0.0
        assert e["pos/dx"] <= 1e-3 and e["Fe"] <= 1e-3, e
    assert rep.step_ms_p50 > 0 and rep.step_ms_p99 >= rep.step_ms_p50


def test_served_nan_scene_is_evicted_others_equal_alone():
    """One request's state gets a NaN velocity (written through
    mpm_write_particles after its build): the batch health pass evicts it
    with an error record, and the requests sharing its batch still equal the
    same requests run alone."""
    from mpm_amd.serve import DeviceBackend
    stream = scenes.request_stream(seed=42, rate=40.0, duration=0.5, n_lo=2000, n_hi=6000, steps_lo=24,
                                   steps_hi=40, max_requests=6)
    reqs = requests_from_scenes(stream)
    # poison a request whose batch has company
    tr = [r.desc.transfer for r in reqs]
    bad = next(r for r in reqs if tr.count(r.desc.transfer) >= 2)
    print("requests:", [(r.rid, r.particles.count, r.steps, r.desc.transfer) for r in reqs], "poisoned:", bad.rid,
          flush=True)

    class Poison(DeviceBackend):
        def create_solver(self, desc, particles):
            s = super().create_solver(desc, particles)
            if particles is bad.particles:
                st = s.read_particles()
                st.velocity[7, 1] = np.nan
                s.write_particles(st)
            return s

    srv = Server(tick_steps=8, keep_results=True, backend=Poison())
    try:
        rep = srv.run(reqs)
    finally:
        srv.close()
    print("serve report:", rep.summary(), flush=True)
    recs = {r.rid: r for r in rep.records}
    print("poisoned record:", recs[bad.rid].error, flush=True)
    assert "non-finite" in recs[bad.rid].error and recs[bad.rid].result is None
    assert rep.failed == 1 and rep.completed == len(reqs) - 1
    for r in reqs:
        if r.rid == bad.rid:
            continue
        alone = Solver(r.desc, r.particles)
        alone.step(r.steps)
        pa = alone.read_particles()
        alone.close()
        pm = recs[r.rid].result
        dx = r.desc.stride
        e = {"pos/dx": float(np.abs(pm.position.astype(np.float64) - pa.position).max() / dx),
             "Fe": rel_err(pm.Fe, pa.Fe)}
        report(f"beside a NaN scene: rid={r.rid} N={r.particles.count}", e)
        assert np.all(np.isfinite(pm.position)) and e["pos/dx"] <= 1e-3 and e["Fe"] <= 1e-3, e


def test_command_line_driver(capsys):
    from mpm_amd import serve

    assert serve.main(["--seconds", "2", "--rate", "4", "--max-requests", "4", "--burst"]) == 0
    rep = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    print("driver report:", rep, flush=True)
    assert rep["completed"] == rep["requests"] > 0 and rep["failed"] == 0 and rep["errors"] == []
