"""Host logic of the continuous-batching server (mpm_amd.serve) with a fake
device backend and a fake clock: admission by HBM budget in arrival order,
joins while other scenes run, exact step counts, budget accounting."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.serve import Server, ServeRequest, estimate_bytes


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeSolver:
    def __init__(self, desc, particles, log):
        self.desc, self.count, self.steps, self.closed = desc, particles.count, 0, False
        self.log = log

    def read_particles(self):
        return {"steps": self.steps}

    def close(self):
        self.closed = True


class FakeBatch:
    def __init__(self, key, clock, log):
        self.key, self.clock, self.log, self.members = key, clock, log, []

    def add(self, s):
        assert s.desc.interp == self.key[0] and s.desc.transfer == self.key[1]
        self.members.append(s)
        self.log.append(("add", self.clock.t, id(s)))

    def remove(self, s):
        self.members.remove(s)

    def step(self, k):
        for m in self.members:
            m.steps += k
        self.clock.t += 1e-3 * k * (1 + sum(m.count for m in self.members) / 1e5)

    def synchronize(self):
        pass

    def read_times(self):
        return []

    def close(self):
        pass


class FakeBackend:
    def __init__(self, clock):
        self.clock, self.log, self.solvers = clock, [], []

    def create_solver(self, desc, particles):
        self.clock.t += 2e-3  # build cost
        s = FakeSolver(desc, particles, self.log)
        self.solvers.append(s)
        return s

    def create_batch(self, interp, transfer):
        return FakeBatch((interp, transfer), self.clock, self.log)


def _req(rid, arrival, n, steps, transfer=N.APIC, res=32):
    desc = scenes.SolverDesc((0, 0, 0), (res, res, res), 1.0 / res, 0.9, 0.95, N.QUADRATIC_BSPLINE, transfer, 1e-4,
                             9.8)
    pos = np.full((n, 3), 0.5, np.float32)
    return ServeRequest(rid, arrival, desc, scenes.snow_particles(pos, (0, 0, 0), 1e-4), steps)


def _server(budget, **kw):
    c = Clock()
    be = FakeBackend(c)
    return Server(budget_bytes=budget, backend=be, clock=c, sleep=c.sleep, keep_results=True, **kw), be, c


def test_every_request_gets_exactly_its_steps():
    reqs = [_req(i, 0.01 * i, 1000 + 100 * i, 37 + 13 * i, N.APIC if i % 2 else N.PIC_FLIP) for i in range(8)]
    srv, be, _ = _server(1e12, tick_steps=16)
    rep = srv.run(reqs)
    assert rep.completed == 8
    for r in rep.records:
        assert r.result == {"steps": r.steps}
    assert all(s.closed for s in be.solvers)
    assert srv.used == 0
    assert set(srv.batches) == {(N.QUADRATIC_BSPLINE, N.APIC), (N.QUADRATIC_BSPLINE, N.PIC_FLIP)}
    assert rep.particle_updates == sum(r.particles * r.steps for r in rep.records)


def test_budget_limits_resident_scenes_and_keeps_arrival_order():
    reqs = [_req(i, 0.0, 2000, 50) for i in range(6)]
    one = estimate_bytes(reqs[0].desc, 2000)
    srv, _, _ = _server(2.5 * one)
    rep = srv.run(reqs)
    assert rep.peak_scenes == 2 and rep.peak_bytes <= 2.5 * one
    adm = [r.admitted for r in sorted(rep.records, key=lambda r: r.rid)]
    assert adm == sorted(adm)
    assert rep.completed == 6 and srv.used == 0


def test_late_arrival_joins_running_batch():
    reqs = [_req(0, 0.0, 5000, 400), _req(1, 0.05, 5000, 40)]
    srv, _, _ = _server(1e12, tick_steps=8)
    rep = srv.run(reqs)
    r0, r1 = sorted(rep.records, key=lambda r: r.rid)
    # request 1 is admitted and finishes while request 0 is still running
    assert r0.admitted <= r1.admitted < r1.finished < r0.finished
    assert r1.admitted - r1.arrival < 0.05


def test_oversized_request_is_rejected():
    big = _req(0, 0.0, 100000, 10)
    srv, _, _ = _server(estimate_bytes(big.desc, 100000) - 1)
    with pytest.raises(ValueError):
        srv.run([big])


def test_estimate_grows_with_particles_and_grid():
    a = _req(0, 0, 1000, 1, res=32).desc
    b = _req(0, 0, 1000, 1, res=64).desc
    assert estimate_bytes(a, 2000) > estimate_bytes(a, 1000)
    assert estimate_bytes(b, 1000) > estimate_bytes(a, 1000)
    assert estimate_bytes(a, 1000) > estimate_bytes(_req(0, 0, 1000, 1, N.PIC_FLIP).desc, 1000)


def test_request_stream_generator_is_seeded():
    a = scenes.request_stream(seed=4, max_requests=5, n_lo=1000, n_hi=2000)
    b = scenes.request_stream(seed=4, max_requests=5, n_lo=1000, n_hi=2000)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.array_equal(x.scene.particles.position, y.scene.particles.position) for x, y in zip(a, b))
    arr = np.array([r.arrival for r in a])
    assert (np.diff(arr) > 0).all()


def test_scene_built_during_the_last_tick_still_runs():
    # request 1 arrives while request 0's only tick is in flight; request 0
    # retires at the end of that tick, leaving only the staged scene
    reqs = [_req(0, 0.0, 1000, 8), _req(1, 0.001, 1000, 5)]
    srv, _, _ = _server(1e12, tick_steps=8)
    rep = srv.run(reqs)
    assert rep.completed == 2
    assert sorted(r.result["steps"] for r in rep.records) == [5, 8]


def test_batch_key_space_limits_admission():
    from mpm_amd.serve import key_span
    big = _req(0, 0.0, 1000, 10, res=1024)
    span = key_span(big.desc)
    assert span == (257 ** 3) * 64 + 64
    reqs = [_req(i, 0.0, 1000, 10, res=1024) for i in range(5)]  # 3 fit in 2^32 keys
    srv, _, _ = _server(1e15)
    rep = srv.run(reqs)
    assert rep.completed == 5 and rep.peak_scenes == (1 << 32) // span
