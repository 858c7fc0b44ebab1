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

Failures stay with the request that caused them.  After every tick the
batch's health pass (one launch) flags members with a non-finite position
or velocity, particles outside their grid, or a CFL number above the limit;
such a member is evicted with an error record and the others continue.  A
library error in a batch's step evicts that batch's members and the other
batches continue; a failed build rejects its request.  Whatever ends the
loop, every scene and batch is freed.

Admission counts device memory as measured where it exists: the handles'
own bytes after their build, the batches' buffers (which grow by 1.5x and
are freed when a batch empties), and the estimate of scenes built but not
yet joined.

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
from concurrent.futures import ThreadPoolExecutor
from concurrent.futures import wait as wait_futures
from dataclasses import dataclass, field

import numpy as np

from . import native as N
from .solver import ParticleInput, SolverDesc

_STAGE_FLOATS_APIC = 43  # x3 v3 B9 Fe9 Fp9 m V0 mu0 lam0 tau6
_STAGE_FLOATS_PIC = 34


# Per-particle words of a batch's buffers: sort keys and values in and out
# (6), flags and item scenes (2), item records (2), store-reorder scratch (9).
_BATCH_WORDS = 19
_BATCH_GROWTH = 1.5  # the batch's buffers grow by this factor and never shrink while it has members


def estimate_bytes(desc: SolverDesc, n: int) -> int:
    """Peak device bytes of one scene: particle store (twice, for the build's
    physical reorder), its sort buffers, its store-reorder scratch (9 words
    per particle), the 10-plane block-major grid with its touched flags, and
    its share of the batch's buffers with their growth slack."""
    f = _STAGE_FLOATS_APIC if desc.transfer == N.APIC else _STAGE_FLOATS_PIC
    own = 2 * (4 * f + 6) + 8 * 4 + 9 * 4 + 4
    share = _BATCH_GROWTH * 4 * _BATCH_WORDS
    blocks = 1
    for r in desc.resolution:
        blocks *= (int(r) + 3) // 4
    grid = blocks * 64 * 10 * 4 + blocks * 4
    return int(int(n) * (own + share) + grid + (1 << 20))


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
    error: str = ""  # why the request was evicted or rejected ("" when it completed)
    steps_done: int = 0


@dataclass
class _Active:
    solver: object
    remaining: int
    rec: RequestRecord
    key: tuple
    span: int = 0
    bytes: int = 0  # the handle's own device bytes, measured after its build


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
    failed: int = 0  # requests evicted or rejected with an error
    records: list = field(default_factory=list, repr=False)

    def summary(self) -> dict:
        d = {k: v for k, v in self.__dict__.items() if k != "records"}
        d["errors"] = [{"rid": r.rid, "particles": r.particles, "steps": r.steps, "steps_done": r.steps_done,
                        "error": r.error} for r in self.records if r.error][:20]
        return d


class Server:
    def __init__(self, budget_bytes: float = 200e9, tick_steps: int = 16, max_scenes: int | None = None,
                 backend=None, clock=time.perf_counter, keep_results: bool = False, sleep=time.sleep,
                 cfl_limit: float = 1.0, evict_escaped: bool = True, builds_per_tick: int = 8,
                 progress_s: float = 0.0, log=print, build_threads: int = 1):
        self.budget = int(budget_bytes)
        # scenes are built on build_threads background threads (0: inline,
        # between ticks): the library's calls on different handles may
        # overlap and ctypes releases the GIL during them, so the main thread
        # keeps the batches stepping while scenes upload and initialise
        self.build_threads = max(0, int(build_threads))
        self._pool = None
        self.building: list = []  # (future, request, record) in submission order
        # a progress line (arrived, resident, done, failed, particle-updates/s) every progress_s seconds
        self.progress_s = float(progress_s)
        self.log = log
        # scenes built per admission pass while others run: a burst of
        # arrivals is built a few at a time between and during ticks, so the
        # GPU steps the first ones while the host builds the rest
        self.builds_per_tick = max(1, int(builds_per_tick))
        self.tick_steps = int(tick_steps)
        self.max_scenes = max_scenes
        self.backend = backend or DeviceBackend()
        self.clock = clock
        self.sleep = sleep
        self.keep_results = keep_results
        self.cfl_limit = float(cfl_limit)
        self.evict_escaped = bool(evict_escaped)
        self.batches: dict[tuple, object] = {}
        self.batch_bytes: dict[tuple, int] = {}
        self.active: dict[int, _Active] = {}
        self.staged: list[_Active] = []
        self.records: list[RequestRecord] = []
        self.step_ms: list[float] = []
        self.tick_step_ms: list[float] = []
        self.peak_bytes = 0
        self.peak_scenes = 0

    # -- device memory
    @property
    def used(self) -> int:
        """Device bytes committed: measured bytes of the joined handles and
        of the batches, estimates of the scenes built but not yet joined."""
        return (sum(a.bytes for a in self.active.values()) + sum(self.batch_bytes.values()) +
                sum(a.rec.est_bytes for a in self.staged) + sum(rec.est_bytes for _, _, rec in self.building))

    def _note_peak(self):
        self.peak_bytes = max(self.peak_bytes, self.used)

    def _measure_batch(self, key):
        b = self.batches.get(key)
        fn = getattr(b, "device_bytes", None)
        if b is not None and fn is not None:
            self.batch_bytes[key] = int(fn())

    # -- admission
    def _fits(self, req: ServeRequest, rec: RequestRecord) -> bool:
        resident = len(self.active) + len(self.staged) + len(self.building)
        if self.max_scenes is not None and resident >= self.max_scenes:
            return False
        key = (req.desc.interp, req.desc.transfer)
        keys = sum(a.span for a in list(self.active.values()) + self.staged if a.key == key)
        keys += sum(key_span(r.desc) for _, r, _ in self.building if (r.desc.interp, r.desc.transfer) == key)
        if keys + key_span(req.desc) > (1 << 32):
            return False
        return self.used + rec.est_bytes <= self.budget

    def _create(self, req: ServeRequest):
        """The scene on the device (its own stream; safe while batch steps
        are in flight, and on a builder thread) and its measured bytes."""
        s = self.backend.create_solver(req.desc, req.particles)
        stats = getattr(s, "stats", None)
        return s, (int(stats()["device_bytes"]) if stats is not None else None)

    def _stage(self, req: ServeRequest, rec: RequestRecord, s, nbytes):
        a = _Active(s, int(req.steps), rec, (req.desc.interp, req.desc.transfer), key_span(req.desc))
        # measured where the backend reports it; else the estimate stands
        a.bytes = nbytes if nbytes is not None else rec.est_bytes
        self.staged.append(a)
        self._note_peak()

    def _build(self, req: ServeRequest, rec: RequestRecord, t0: float):
        """Builds the scene, inline or on the builder thread.  A library
        error rejects the request."""
        rec.admitted = self.clock() - t0
        if self.build_threads > 0:
            if self._pool is None:
                self._pool = ThreadPoolExecutor(self.build_threads, thread_name_prefix="mpm-build")
            self.building.append((self._pool.submit(self._create, req), req, rec))
            self._note_peak()
            return
        try:
            s, nbytes = self._create(req)
        except N.MpmError as e:
            rec.error = f"build failed: {e}"
            rec.finished = self.clock() - t0
            return
        self._stage(req, rec, s, nbytes)

    def _collect(self, t0: float, wait: bool = False):
        """Scenes finished on the builder thread are staged (in submission
        order); wait: block until the first one is done."""
        if wait and self.building:
            wait_futures([self.building[0][0]], timeout=0.05)
        while self.building and self.building[0][0].done():
            f, req, rec = self.building.pop(0)
            e = f.exception()
            if isinstance(e, N.MpmError):
                rec.error = f"build failed: {e}"
                rec.finished = self.clock() - t0
                continue
            if e is not None:
                raise e
            s, nbytes = f.result()
            self._stage(req, rec, s, nbytes)

    def _join(self, a: _Active, t0: float):
        """Adds a built scene to its batch (between ticks)."""
        b = self.batches.get(a.key)
        try:
            if b is None:
                b = self.backend.create_batch(*a.key)
                self.batches[a.key] = b
                self.batch_bytes[a.key] = 0
            b.add(a.solver)
        except N.MpmError as e:
            self._drop(a, t0, f"join failed: {e}", joined=False)
            return
        self.active[a.rec.rid] = a
        self._measure_batch(a.key)
        self.peak_scenes = max(self.peak_scenes, len(self.active))
        self._note_peak()

    def _drop(self, a: _Active, t0: float, error: str = "", joined: bool = True):
        """Takes a scene out (finished or failed): result and final CFL read
        if it finished, handle freed, budget returned."""
        b = self.batches.get(a.key)
        if joined and b is not None:
            try:
                b.remove(a.solver)
            except N.MpmError as e:
                error = error or f"remove failed: {e}"
        if not error:
            try:
                if self.keep_results:
                    a.rec.result = a.solver.read_particles()
                stats = getattr(a.solver, "stats", None)
                if stats is not None:
                    a.rec.cfl = float(stats()["cfl"])
            except N.MpmError as e:
                error = f"readback failed: {e}"
        a.rec.error = error
        try:
            a.solver.close()
        except N.MpmError:
            pass
        a.rec.finished = self.clock() - t0
        self.active.pop(a.rec.rid, None)
        if b is not None and not any(x.key == a.key for x in self.active.values()):
            self._close_batch(a.key)  # an empty batch gives its buffers back
        else:
            self._measure_batch(a.key)

    def _close_batch(self, key):
        b = self.batches.pop(key, None)
        self.batch_bytes.pop(key, None)
        if b is not None:
            try:
                b.close()
            except N.MpmError:
                pass

    def _health(self, b, members, t0):
        """Evicts members whose state is non-finite, left the grid or runs
        above the CFL limit."""
        fn = getattr(b, "health", None)
        if fn is None:
            return
        order = {id(s): k for k, s in enumerate(getattr(b, "members", []))}
        h = fn()
        for a in members:
            k = order.get(id(a.solver))
            if k is None:
                continue
            r = h[k]
            why = ""
            if r["nonfinite"] > 0:
                why = f"{r['nonfinite']} particles with a non-finite position or velocity"
            elif self.evict_escaped and r["escaped"] > 0:
                why = f"{r['escaped']} particles outside the grid"
            elif not (r["cfl"] <= self.cfl_limit):
                why = f"CFL number {r['cfl']:.3g} above the limit {self.cfl_limit:g}"
            if why:
                a.rec.cfl = float(r["cfl"])
                self._drop(a, t0, f"evicted after {a.rec.steps_done} steps: {why}")

    def _progress(self, t0):
        now = self.clock() - t0
        done = [r for r in self.records if r.finished >= 0 and not r.error]
        pu = sum(r.particles * r.steps for r in done)
        self.log(f"[serve] t={now:.0f}s arrived {len(self.records)} resident {len(self.active)} "
                 f"done {len(done)} failed {sum(1 for r in self.records if r.error)} "
                 f"{pu / max(now, 1e-9):.3e} particle-updates/s, {self.used / 1e9:.1f} GB committed")

    # -- main loop
    def run(self, requests) -> ServeReport:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in reqs:
            if estimate_bytes(r.desc, r.particles.count) > self.budget:
                raise ValueError(f"request {r.rid} alone exceeds the HBM budget")
            if key_span(r.desc) > (1 << 32):
                raise ValueError(f"request {r.rid}: grid too large for 32-bit cell keys")
        try:
            return self._run(reqs)
        except BaseException:
            self.close()  # nothing stays allocated when the loop ends by an exception
            raise

    def _run(self, reqs) -> ServeReport:
        pending = deque(reqs)
        queue: deque = deque()
        t0 = self.clock()

        def arrivals():
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                rec = RequestRecord(r.rid, r.particles.count, int(r.steps), r.desc.interp, r.desc.transfer, r.arrival,
                                    estimate_bytes(r.desc, r.particles.count))
                self.records.append(rec)
                queue.append((r, rec))

        def admissions():
            built = 0
            while queue and self._fits(queue[0][0], queue[0][1]):
                if self.build_threads == 0 and built >= self.builds_per_tick:  # the next tick runs what is built
                    break
                r, rec = queue.popleft()
                self._build(r, rec, t0)
                built += 1

        def joins():
            self._collect(t0)
            staged, self.staged = self.staged, []
            for a in staged:
                self._join(a, t0)

        next_report = self.progress_s
        while pending or queue or self.staged or self.active or self.building:
            if self.progress_s > 0 and self.clock() - t0 >= next_report:
                next_report += self.progress_s
                self._progress(t0)
            arrivals()
            admissions()
            joins()
            if not self.active:
                if self.building:
                    self._collect(t0, wait=True)
                elif pending and not queue and not self.staged:
                    self.sleep(max(0.0, min(0.05, pending[0].arrival - (self.clock() - t0))))
                continue
            # one tick: enqueue k steps on every batch, admit meanwhile, then retire
            ticks = []
            tick_start = self.clock()
            for key, b in list(self.batches.items()):
                members = [a for a in self.active.values() if a.key == key]
                if not members:
                    continue
                k = min(self.tick_steps, min(a.remaining for a in members))
                try:
                    b.step(k)
                except N.MpmError as e:
                    for a in members:
                        self._drop(a, t0, f"batch step failed: {e}")
                    continue
                ticks.append((key, b, members, k))
            arrivals()
            admissions()  # builds overlap the enqueued steps; they join after the tick
            done = []
            for key, b, members, k in ticks:
                try:
                    b.synchronize()
                except N.MpmError as e:
                    for a in members:
                        self._drop(a, t0, f"batch step failed: {e}")
                    continue
                for a in members:
                    a.remaining -= k
                    a.rec.steps_done += k
                done.append((key, b, members, k))
            tick_wall = (self.clock() - tick_start) * 1e3
            kmax = max((k for *_, k in ticks), default=0)
            if kmax > 0:
                self.tick_step_ms.extend([tick_wall / kmax] * kmax)
            for key, b, members, k in done:
                try:
                    if hasattr(b, "read_times"):
                        self.step_ms.extend(float(x) for x in b.read_times())
                    self._health(b, [a for a in members if a.rec.rid in self.active], t0)
                except N.MpmError as e:
                    for a in [a for a in members if a.rec.rid in self.active]:
                        self._drop(a, t0, f"health check failed: {e}")
            for a in [a for a in self.active.values() if a.remaining <= 0]:
                self._drop(a, t0)
        wall = self.clock() - t0
        done = [r for r in self.records if r.finished >= 0 and not r.error]
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
            max_cfl=max((r.cfl for r in done), default=0.0),
            failed=sum(1 for r in self.records if r.error), records=self.records)

    def close(self):
        for f, _, _ in self.building:  # scenes still being built are freed once built
            try:
                f.result()[0].close()
            except Exception:
                pass
        self.building = []
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None
        for a in list(self.active.values()) + self.staged:
            try:
                a.solver.close()
            except N.MpmError:
                pass
        self.active.clear()
        self.staged = []
        for key in list(self.batches):
            self._close_batch(key)


def requests_from_scenes(stream) -> list[ServeRequest]:
    """ServeRequests from scenes.request_stream() output."""
    return [ServeRequest(i, r.arrival, r.scene.desc, r.scene.particles, r.scene.steps) for i, r in enumerate(stream)]


def main(argv=None) -> int:
    """Serve a seeded request stream and print the report as one JSON line:
    ``python3 -m mpm_amd.serve --seconds 60 --rate 20`` (``--burst``: every
    request arrives at t=0, the server's capacity)."""
    import argparse
    import dataclasses
    import json

    from . import scenes

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second")
    ap.add_argument("--seconds", type=float, default=60.0, help="arrival window")
    ap.add_argument("--max-requests", type=int, default=None)
    ap.add_argument("--burst", action="store_true", help="all requests arrive at t=0")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--tick-steps", type=int, default=16)
    ap.add_argument("--progress", type=float, default=30.0, help="seconds between progress lines (0: none)")
    a = ap.parse_args(argv)
    stream = []
    for r in scenes.iter_request_stream(seed=a.seed, rate=a.rate, duration=a.seconds, max_requests=a.max_requests):
        stream.append(r)
        if len(stream) % 1000 == 0:
            print(f"[serve] generated {len(stream)} requests", flush=True)
    reqs = requests_from_scenes(stream)
    if a.burst:
        reqs = [dataclasses.replace(r, arrival=0.0) for r in reqs]
    print(f"[serve] {len(reqs)} requests", flush=True)
    srv = Server(budget_bytes=a.budget_gb * 1e9, tick_steps=a.tick_steps, progress_s=a.progress,
                 log=lambda m: print(m, flush=True))
    try:
        rep = srv.run(reqs)
    finally:
        srv.close()
    print(json.dumps(rep.summary()), flush=True)
    return 0  # evicted or rejected requests are in the report ("failed", "errors")


if __name__ == "__main__":
    raise SystemExit(main())
