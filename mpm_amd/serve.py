"""Continuous-batching request loop over device-resident scenes.

Requests (a scene plus a step count) arrive over time.  The server admits
them in arrival order while their estimated device memory fits the HBM
budget, builds each admitted scene on the device (upload + initialize, on
the scene's own stream, overlapping the batches already running), and adds
it to the batch of its (interpolation, transfer) pair.  Every tick it
enqueues k steps on each batch (k = the smallest remaining step count of
the batch's members, capped at ``tick_steps``, so nobody overshoots), then
retires the members that are done: their state is read back, their memory
freed and their budget returned.  Scenes stay on the device from admission
to completion; requests that arrive while others run join at the next tick.

Reported: per-step latency (device time of each batch step, hipEvents, and
host wall time per step of a tick), per-request latency (finish - arrival)
and queueing delay as p50/p99, and aggregate particle-updates/s.

The reference has no serving layer (one scene per process, driven by its
demo loop, src/demo/mpm-apic/main.cu:47-209); this is the MI355X
addition for the request-stream workload.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import native as N
from .solver import ParticleInput, SolverDesc

_STAGE_FLOATS_APIC = 43  # x3 v3 B9 Fe9 Fp9 m V0 mu0 lam0 tau6
_STAGE_FLOATS_PIC = 34


def estimate_bytes(desc: SolverDesc, n: int) -> int:
    """Peak device bytes of one scene: particle store (twice, for the build's
    physical reorder), its sort buffers and the batch's per-particle share,
    the store-reorder scratch (9 words per particle, the scene's and its
    batch's), the 10-plane block-major grid and its touched flags."""
    f = _STAGE_FLOATS_APIC if desc.transfer == N.APIC else _STAGE_FLOATS_PIC
    per_particle = 2 * (4 * f + 6) + 8 * 4 + 9 * 4 + 4 + 2 * 9 * 4
    blocks = 1
    for r in desc.resolution:
        blocks *= (int(r) + 3) // 4
    grid = blocks * 64 * 10 * 4 + blocks * 4
    return int(n) * per_particle + grid + (1 << 20)


def key_span(desc: SolverDesc) -> int:
    """Keys a scene occupies in its batch's 32-bit key space: its cell keys
    (64 per 4x4x4 key block, ((res+3)>>2)+1 blocks per axis) plus the
    escape block."""
    k = 1
    for r in desc.resolution:
        k *= ((int(r) + 3) >> 2) + 1
    return k * 64 + 64


@dataclass
class ServeRequest:
    rid: int
    arrival: float  # seconds after the server starts
    desc: SolverDesc
    particles: ParticleInput
    steps: int


@dataclass
class RequestRecord:
    rid: int
    particles: int
    steps: int
    interp: int
    transfer: int
    arrival: float
    est_bytes: int
    admitted: float = -1.0
    finished: float = -1.0
    cfl: float = 0.0  # max|v| dt / h of the final state (CFL monitor)
    result: object = None


@dataclass
class _Active:
    solver: object
    remaining: int
    rec: RequestRecord
    key: tuple
    span: int = 0


class DeviceBackend:
    """Scenes and batches on the GPU (libmpm_hip.so)."""

    def create_solver(self, desc, particles):
        from .solver import Solver
        return Solver(desc, particles)

    def create_batch(self, interp, transfer):
        from .batch import Batch
        b = Batch(interp, transfer)
        b.timing(True)
        return b


def _pct(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else 0.0


@dataclass
class ServeReport:
    requests: int
    completed: int
    wall_s: float
    particle_updates: int
    particle_updates_per_s: float
    step_ms_p50: float
    step_ms_p99: float
    tick_step_ms_p50: float
    tick_step_ms_p99: float
    request_s_p50: float
    request_s_p99: float
    queue_s_p50: float
    queue_s_p99: float
    peak_bytes: int
    peak_scenes: int
    max_cfl: float = 0.0  # largest final-state CFL number over the requests
    records: list = field(default_factory=list, repr=False)

    def summary(self) -> dict:
        return {k: v for k, v in self.__dict__.items() if k != "records"}


class Server:
    def __init__(self, budget_bytes: float = 200e9, tick_steps: int = 16, max_scenes: int | None = None,
                 backend=None, clock=time.perf_counter, keep_results: bool = False, sleep=time.sleep):
        self.budget = int(budget_bytes)
        self.tick_steps = int(tick_steps)
        self.max_scenes = max_scenes
        self.backend = backend or DeviceBackend()
        self.clock = clock
        self.sleep = sleep
        self.keep_results = keep_results
        self.used = 0
        self.batches: dict[tuple, object] = {}
        self.active: dict[int, _Active] = {}
        self.records: list[RequestRecord] = []
        self.step_ms: list[float] = []
        self.tick_step_ms: list[float] = []
        self.peak_bytes = 0
        self.peak_scenes = 0

    # -- admission
    def _fits(self, req: ServeRequest, rec: RequestRecord, staged: list) -> bool:
        if self.max_scenes is not None and len(self.active) + len(staged) >= self.max_scenes:
            return False
        key = (req.desc.interp, req.desc.transfer)
        keys = sum(a.span for a in list(self.active.values()) + staged if a.key == key)
        if keys + key_span(req.desc) > (1 << 32):
            return False
        return self.used + rec.est_bytes <= self.budget

    def _build(self, req: ServeRequest, rec: RequestRecord, t0: float) -> _Active:
        """Creates the scene on the device (its own stream): safe while
        batch steps are in flight."""
        s = self.backend.create_solver(req.desc, req.particles)
        self.used += rec.est_bytes
        rec.admitted = self.clock() - t0
        self.peak_bytes = max(self.peak_bytes, self.used)
        return _Active(s, int(req.steps), rec, (req.desc.interp, req.desc.transfer), key_span(req.desc))

    def _join(self, a: _Active):
        """Adds a built scene to its batch (between ticks)."""
        b = self.batches.get(a.key)
        if b is None:
            b = self.backend.create_batch(*a.key)
            self.batches[a.key] = b
        b.add(a.solver)
        self.active[a.rec.rid] = a
        self.peak_scenes = max(self.peak_scenes, len(self.active))

    def _retire(self, a: _Active, t0: float):
        b = self.batches[a.key]
        b.remove(a.solver)
        if self.keep_results:
            a.rec.result = a.solver.read_particles()
        stats = getattr(a.solver, "stats", None)
        if stats is not None:
            a.rec.cfl = float(stats()["cfl"])
        a.solver.close()
        self.used -= a.rec.est_bytes
        a.rec.finished = self.clock() - t0
        del self.active[a.rec.rid]

    # -- main loop
    def run(self, requests) -> ServeReport:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in reqs:
            if estimate_bytes(r.desc, r.particles.count) > self.budget:
                raise ValueError(f"request {r.rid} alone exceeds the HBM budget")
            if key_span(r.desc) > (1 << 32):
                raise ValueError(f"request {r.rid}: grid too large for 32-bit cell keys")
        pending = deque(reqs)
        queue: deque = deque()
        recs = {}
        t0 = self.clock()

        def arrivals():
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                rec = RequestRecord(r.rid, r.particles.count, int(r.steps), r.desc.interp, r.desc.transfer, r.arrival,
                                    estimate_bytes(r.desc, r.particles.count))
                recs[r.rid] = rec
                self.records.append(rec)
                queue.append((r, rec))

        staged: list[_Active] = []

        def admissions():
            while queue and self._fits(queue[0][0], queue[0][1], staged):
                r, rec = queue.popleft()
                staged.append(self._build(r, rec, t0))

        def joins():
            for a in staged:
                self._join(a)
            staged.clear()

        while pending or queue or staged or self.active:
            arrivals()
            admissions()
            joins()
            if not self.active:
                if pending:
                    self.sleep(max(0.0, min(0.05, pending[0].arrival - (self.clock() - t0))))
                continue
            # one tick: enqueue k steps on every batch, admit meanwhile, then retire
            ticks = []
            tick_start = self.clock()
            for key, b in self.batches.items():
                members = [a for a in self.active.values() if a.key == key]
                if not members:
                    continue
                k = min(self.tick_steps, min(a.remaining for a in members))
                b.step(k)
                ticks.append((b, members, k))
            arrivals()
            admissions()  # builds overlap the enqueued steps; they join after the tick
            for b, members, k in ticks:
                b.synchronize()
                for a in members:
                    a.remaining -= k
            tick_wall = (self.clock() - tick_start) * 1e3
            kmax = max(k for _, _, k in ticks)
            if kmax > 0:
                self.tick_step_ms.extend([tick_wall / kmax] * kmax)
            for b, _, _ in ticks:
                if hasattr(b, "read_times"):
                    self.step_ms.extend(float(x) for x in b.read_times())
            for a in [a for a in self.active.values() if a.remaining <= 0]:
                self._retire(a, t0)
        wall = self.clock() - t0
        done = [r for r in self.records if r.finished >= 0]
        pu = sum(r.particles * r.steps for r in done)
        return ServeReport(
            requests=len(self.records), completed=len(done), wall_s=wall, particle_updates=pu,
            particle_updates_per_s=pu / wall if wall > 0 else 0.0,
            step_ms_p50=_pct(self.step_ms, 50), step_ms_p99=_pct(self.step_ms, 99),
            tick_step_ms_p50=_pct(self.tick_step_ms, 50), tick_step_ms_p99=_pct(self.tick_step_ms, 99),
            request_s_p50=_pct([r.finished - r.arrival for r in done], 50),
            request_s_p99=_pct([r.finished - r.arrival for r in done], 99),
            queue_s_p50=_pct([r.admitted - r.arrival for r in done], 50),
            queue_s_p99=_pct([r.admitted - r.arrival for r in done], 99),
            peak_bytes=self.peak_bytes, peak_scenes=self.peak_scenes,
            max_cfl=max((r.cfl for r in done), default=0.0), records=self.records)

    def close(self):
        for a in list(self.active.values()):
            a.solver.close()
        self.active.clear()
        for b in self.batches.values():
            b.close()
        self.batches.clear()


def requests_from_scenes(stream) -> list[ServeRequest]:
    """ServeRequests from scenes.request_stream() output."""
    return [ServeRequest(i, r.arrival, r.scene.desc, r.scene.particles, r.scene.steps) for i, r in enumerate(stream)]
