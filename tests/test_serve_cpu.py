"""Host logic of the continuous-batching server (mpm_amd.serve) with a fake
device backend and a fake clock: admission by HBM budget in arrival order,
joins while other scenes run, exact step counts, budget accounting."""
from __future__ import annotations

import time

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
        self.clock, self.log, self.solvers, self.batch_keys = clock, [], [], []

    def create_solver(self, desc, particles):
        self.clock.t += 2e-3  # build cost
        s = FakeSolver(desc, particles, self.log)
        self.solvers.append(s)
        return s

    def create_batch(self, interp, transfer):
        self.batch_keys.append((interp, transfer))
        return FakeBatch((interp, transfer), self.clock, self.log)


def _req(rid, arrival, n, steps, transfer=N.APIC, res=32):
    desc = scenes.SolverDesc((0, 0, 0), (res, res, res), 1.0 / res, 0.9, 0.95, N.QUADRATIC_BSPLINE, transfer, 1e-4,
                             9.8)
    pos = np.full((n, 3), 0.5, np.float32)
    return ServeRequest(rid, arrival, desc, scenes.snow_particles(pos, (0, 0, 0), 1e-4), steps)


def _server(budget, **kw):
    c = Clock()
    be = FakeBackend(c)
    kw.setdefault("build_threads", 0)  # inline builds: the fake clock stays deterministic
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
    assert set(be.batch_keys) == {(N.QUADRATIC_BSPLINE, N.APIC), (N.QUADRATIC_BSPLINE, N.PIC_FLIP)}
    assert srv.batches == {}  # emptied batches are freed
    assert rep.particle_updates == sum(r.particles * r.steps for r in rep.records)


def test_burst_is_built_a_few_scenes_per_tick():
    """20 requests arriving at once: the first tick runs after 4 builds, and
    the rest are built between and during ticks, still in arrival order."""
    reqs = [_req(i, 0.0, 2000, 300) for i in range(20)]
    srv, be, _ = _server(1e12, tick_steps=16, builds_per_tick=4)
    events = []
    create = be.create_solver

    def logged_create(desc, particles):
        events.append("build")
        return create(desc, particles)

    be.create_solver = logged_create
    step = FakeBatch.step

    def logged_step(self, k):
        events.append("step")
        step(self, k)

    FakeBatch.step = logged_step
    try:
        rep = srv.run(reqs)
    finally:
        FakeBatch.step = step
    assert events[:5] == ["build"] * 4 + ["step"], events[:8]
    assert events.count("build") == 20 and rep.completed == 20
    adm = [r.admitted for r in sorted(rep.records, key=lambda r: r.rid)]
    assert adm == sorted(adm)
    for r in rep.records:
        assert r.result == {"steps": r.steps}


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


# ------------------------------------------------------ failure isolation
class HealthBatch(FakeBatch):
    """Fake batch with a health pass, measured buffer bytes that grow by 1.5x
    and never shrink (as the device batch's), and planted failures: members
    named in `nan_after` report non-finite particles once they have taken
    that many steps; a step raises MpmError when `fail_steps` is set."""

    PER_PARTICLE = 4 * 19  # batch words per particle, as the device batch

    def __init__(self, key, clock, log, nan_after, fail_steps, meter):
        super().__init__(key, clock, log)
        self.nan_after, self.fail_steps, self.meter, self.cap = nan_after, fail_steps, meter, 0

    def add(self, s):
        super().add(s)
        n = sum(m.count for m in self.members)
        if n > self.cap:
            self.cap = max(n, int(self.cap * 1.5))
        self.meter.check()

    def step(self, k):
        if self.fail_steps:
            raise N.MpmError(N.MPM_ERR_HIP, "planted batch failure")
        super().step(k)

    def device_bytes(self):
        return self.cap * self.PER_PARTICLE

    def health(self):
        out = []
        for m in self.members:
            bad = m.rid in self.nan_after and m.steps >= self.nan_after[m.rid]
            out.append({"max_speed": float("nan") if bad else 1.0, "cfl": float("nan") if bad else 0.01,
                        "nonfinite": 3 if bad else 0, "escaped": 0})
        return out

    def close(self):
        self.closed = True
        self.meter.batches.remove(self)


class Meter:
    """Real device bytes of the fake backend at every change, against the budget."""

    def __init__(self, budget):
        self.budget, self.solvers, self.batches, self.peak = budget, [], [], 0

    def check(self):
        now = sum(s.bytes for s in self.solvers if not s.closed) + sum(b.device_bytes() for b in self.batches)
        self.peak = max(self.peak, now)
        assert now <= self.budget, (now, self.budget)


class MeteredSolver(FakeSolver):
    def __init__(self, desc, particles, log, rid, meter):
        super().__init__(desc, particles, log)
        self.rid, self.meter = rid, meter
        # own bytes: the estimate without the batch share (what the handle holds)
        from mpm_amd.serve import _BATCH_GROWTH, _BATCH_WORDS
        self.bytes = estimate_bytes(desc, self.count) - int(self.count * _BATCH_GROWTH * 4 * _BATCH_WORDS)

    def stats(self):
        return {"device_bytes": self.bytes, "cfl": 0.01}

    def close(self):
        self.closed = True
        self.meter.check()


class FailingBackend(FakeBackend):
    def __init__(self, clock, budget, nan_after=None, fail_keys=(), fail_build=()):
        super().__init__(clock)
        self.meter = Meter(budget)
        self.nan_after, self.fail_keys, self.fail_build = nan_after or {}, set(fail_keys), set(fail_build)
        self.made = []

    def create_solver(self, desc, particles):
        rid = len(self.made)
        self.made.append(rid)
        if rid in self.fail_build:
            raise N.MpmError(N.MPM_ERR_OOM, "planted build failure")
        self.clock.t += 2e-3
        s = MeteredSolver(desc, particles, self.log, rid, self.meter)
        self.solvers.append(s)
        self.meter.solvers.append(s)
        self.meter.check()
        return s

    def create_batch(self, interp, transfer):
        self.batch_keys.append((interp, transfer))
        b = HealthBatch((interp, transfer), self.clock, self.log, self.nan_after, (interp, transfer) in self.fail_keys,
                        self.meter)
        self.meter.batches.append(b)
        return b


def _failing(budget=1e12, **kw):
    c = Clock()
    be = FailingBackend(c, budget, **{k: kw.pop(k) for k in ("nan_after", "fail_keys", "fail_build") if k in kw})
    kw.setdefault("build_threads", 0)
    return Server(budget_bytes=budget, backend=be, clock=c, sleep=c.sleep, keep_results=True, **kw), be


def test_nan_scene_is_evicted_and_others_continue():
    reqs = [_req(i, 0.0, 1000 + 10 * i, 60) for i in range(4)]
    srv, be = _failing(tick_steps=16, nan_after={2: 20})
    rep = srv.run(reqs)
    recs = {r.rid: r for r in rep.records}
    assert "non-finite" in recs[2].error and recs[2].result is None
    assert recs[2].steps_done == 32  # found by the health pass of the tick that produced it
    assert rep.failed == 1 and rep.completed == 3
    for k in (0, 1, 3):
        assert recs[k].error == "" and recs[k].result == {"steps": 60}
    assert all(s.closed for s in be.solvers) and srv.used == 0 and not srv.batches
    assert rep.particle_updates == sum(recs[k].particles * 60 for k in (0, 1, 3))


def test_failing_batch_does_not_stop_the_other():
    reqs = [_req(i, 0.0, 1000, 40, N.APIC if i % 2 else N.PIC_FLIP) for i in range(6)]
    srv, be = _failing(tick_steps=8, fail_keys=[(N.QUADRATIC_BSPLINE, N.APIC)])
    rep = srv.run(reqs)
    for r in rep.records:
        if r.transfer == N.APIC:
            assert "planted batch failure" in r.error
        else:
            assert r.error == "" and r.result == {"steps": 40}
    assert rep.failed == 3 and rep.completed == 3
    assert all(s.closed for s in be.solvers) and not srv.batches


def test_failed_build_rejects_only_that_request():
    reqs = [_req(i, 0.0, 1000, 20) for i in range(3)]
    srv, be = _failing(fail_build=[1])
    rep = srv.run(reqs)
    recs = {r.rid: r for r in rep.records}
    assert "build failed" in recs[1].error
    assert recs[0].result == {"steps": 20} and recs[2].result == {"steps": 20}
    assert rep.failed == 1 and rep.completed == 2


def test_unexpected_error_frees_everything():
    reqs = [_req(i, 0.0, 1000, 50) for i in range(3)]
    srv, be = _failing()

    def broken_read():
        raise RuntimeError("host failure")

    be_create = be.create_solver

    def create(desc, particles):
        s = be_create(desc, particles)
        s.read_particles = broken_read
        return s

    be.create_solver = create
    with pytest.raises(RuntimeError):
        srv.run(reqs)
    assert all(s.closed for s in be.solvers)
    assert not srv.active and not srv.batches and not be.meter.batches


def test_budget_holds_real_bytes_under_churn():
    """Joins and departures of uneven scenes: the device bytes the fake
    backend really holds (handles + batch buffers that only grow) stay within
    the budget at every change."""
    rng = np.random.default_rng(3)
    reqs = [_req(i, 0.002 * i, int(rng.integers(1000, 20000)), int(rng.integers(8, 80))) for i in range(40)]
    budget = 4 * estimate_bytes(reqs[0].desc, 20000)
    srv, be = _failing(budget=budget, tick_steps=8)
    rep = srv.run(reqs)
    assert rep.completed == 40 and rep.failed == 0
    assert be.meter.peak <= budget and rep.peak_bytes <= budget
    assert srv.used == 0


class SlowBuildBackend:
    """Real-time fake device: a build takes 20 ms (off the main thread when
    the server has a builder), a batch step 2 ms."""

    def __init__(self):
        self.log, self.solvers, self.threads = [], [], set()

    def create_solver(self, desc, particles):
        import threading

        self.threads.add(threading.get_ident())
        time.sleep(0.02)
        s = FakeSolver(desc, particles, self.log)
        self.solvers.append(s)
        return s

    def create_batch(self, interp, transfer):
        b = FakeBatch((interp, transfer), Clock(), self.log)
        step = b.step

        def timed_step(k):
            time.sleep(0.002 * k)
            step(k)

        b.step = timed_step
        return b


def test_builder_thread_overlaps_steps():
    """Builds run on the builder thread while the main thread steps: the
    requests complete with exact step counts, in less time than building
    and stepping one after the other would take."""
    import threading

    reqs = [_req(i, 0.0, 2000, 40) for i in range(12)]
    be = SlowBuildBackend()
    srv = Server(budget_bytes=1e12, backend=be, keep_results=True, tick_steps=16, build_threads=1)
    t0 = time.perf_counter()
    rep = srv.run(reqs)
    wall = time.perf_counter() - t0
    srv.close()
    assert rep.completed == 12 and rep.failed == 0
    for r in rep.records:
        assert r.result == {"steps": r.steps}
    assert threading.get_ident() not in be.threads  # every build ran off the main thread
    assert all(s.closed for s in be.solvers)
    print(f"builder thread: 12 builds of 20 ms with 40 steps each in {wall:.3f} s", flush=True)
    assert wall < 12 * 0.02 + 0.25  # the stepping overlapped the builds
