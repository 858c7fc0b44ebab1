"""Seeded synthetic scenes for the benchmark workloads and parity tests.

All randomness comes from a counter-based SplitMix64 stream owned by this
module, so the CPU oracle and the GPU solver receive bit-identical inputs
for a given seed.  Units are SI; the grid origin is 0.

Snow material shared by the workloads (Stomakhin et al. 2013 values as in
the reference configs, reference config/snow_cube.yaml:5-10): E0 = 1.4e5,
nu = 0.2, xi = 10, theta_c = 2.5e-2, theta_s = 7.5e-3, density 400 kg/m^3,
boundary friction 0.9, gravity 9.8.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import native as N
from .solver import ParticleInput, SolverDesc

RHO = 400.0
E0 = 1.4e5
NU = 0.2
XI = 10.0
THETA_C = 2.5e-2
THETA_S = 7.5e-3
FRICTION = 0.9
GRAVITY = 9.8

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """SplitMix64 output for counters 0..n-1 of stream ``seed`` (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counters.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Rng:
    """Counter-based stream: every draw advances a 64-bit counter."""

    def __init__(self, seed: int):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.ctr = 0

    def raw(self, n: int) -> np.ndarray:
        c = np.arange(self.ctr, self.ctr + n, dtype=np.uint64)
        self.ctr += n
        return splitmix64(self.seed, c)

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        u = (self.raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int, sigma=1.0) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return sigma * z

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        return (lo + np.floor(self.uniform(n) * (hi - lo))).astype(np.int64)

    def child(self, k: int) -> "Rng":
        return Rng(int(splitmix64(self.seed ^ 0x5DEECE66D, np.array([k], np.uint64))[0]))


@dataclass
class Scene:
    name: str
    desc: SolverDesc
    particles: ParticleInput
    steps: int
    seed: int
    notes: str = ""


def snow_particles(pos: np.ndarray, vel, mass, young=E0, poisson=NU, hardening=XI, theta_c=THETA_C,
                   theta_s=THETA_S, model=N.FIXED_COROTATED) -> ParticleInput:
    n = pos.shape[0]
    v = np.broadcast_to(np.asarray(vel, dtype=np.float64), (n, 3))
    return ParticleInput(pos, v, mass, hardening, young, poisson, theta_c, theta_s, model)


# ---------------------------------------------------------------- config 0
def snow_cube(seed: int = 0, n: int = 16384, res: int = 64, lo: float = 0.4, hi: float = 0.6,
              transfer: int = N.PIC_FLIP, interp: int = N.QUADRATIC_BSPLINE, dt: float = 1e-4,
              velocity=(0.0, 0.0, 0.0), steps: int = 100) -> Scene:
    """Plain snow cube: n uniform-random particles in [lo,hi]^3 of the unit
    domain, res^3 grid (h = 1/res), mass = rho * box volume / n."""
    r = Rng(seed)
    pos = r.uniform(3 * n, lo, hi).reshape(n, 3)
    mass = RHO * (hi - lo) ** 3 / n
    desc = SolverDesc((0.0, 0.0, 0.0), (res, res, res), 1.0 / res, FRICTION, 0.95, interp, transfer, dt, GRAVITY)
    return Scene("snow_cube", desc, snow_particles(pos, velocity, mass), steps, seed,
                 f"{n} particles U[{lo},{hi}]^3, {res}^3 grid")


# ---------------------------------------------------------------- config 1
def ball_points(rng: Rng, center, radius, h, ppc) -> np.ndarray:
    """Rejection sampling: ppc uniform points in every cell that intersects
    the ball, kept if inside the ball."""
    c = np.asarray(center, np.float64)
    lo = np.floor((c - radius) / h).astype(int)
    hi = np.floor((c + radius) / h).astype(int)
    ii = np.stack(np.meshgrid(*[np.arange(lo[a], hi[a] + 1) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    # cell intersects ball: distance from centre to the cell box <= r
    box_lo = ii * h
    d = np.maximum(np.maximum(box_lo - c, 0.0), c - (box_lo + h))
    ii = ii[(d * d).sum(1) <= radius * radius]
    m = ii.shape[0]
    u = rng.uniform(m * ppc * 3).reshape(m, ppc, 3)
    p = (ii[:, None, :] + u) * h
    p = p.reshape(-1, 3)
    return p[((p - c) ** 2).sum(1) <= radius * radius]


def two_balls(seed: int = 1, res: int = 128, ppc: int = 8, steps: int = 500, dt: float = 5e-5,
              interp: int = N.CUBIC_BSPLINE, transfer: int = N.APIC, radius: float = 0.1, speed: float = 3.0) -> Scene:
    """Two colliding snow balls (r = 0.1 at x = 0.3 and 0.7), +-speed on x."""
    r = Rng(seed)
    h = 1.0 / res
    a = ball_points(r, (0.3, 0.5, 0.5), radius, h, ppc)
    b = ball_points(r, (0.7, 0.5, 0.5), radius, h, ppc)
    mass = RHO * h ** 3 / ppc
    pa = snow_particles(a, (speed, 0, 0), mass)
    pb = snow_particles(b, (-speed, 0, 0), mass)
    desc = SolverDesc((0.0, 0.0, 0.0), (res, res, res), h, FRICTION, 0.95, interp, transfer, dt, GRAVITY)
    return Scene("two_balls", desc, ParticleInput.concat([pa, pb]), steps, seed,
                 f"{a.shape[0]}+{b.shape[0]} particles at {ppc}/cell, {res}^3")


# ---------------------------------------------------------------- config 2
def apic_slab(seed: int = 2, nx: int = 256, ny: int = 128, nz: int = 256, res: int = 256, steps: int = 200,
              dt: float = 2.5e-5, sigma_v: float = 0.5, lo=(0.1, 0.1, 0.1), hi=(0.9, 0.5, 0.9)) -> Scene:
    """Large APIC slab (roofline run): nx*ny*nz jittered lattice in the slab,
    v ~ N(0, sigma_v^2) per component, E = E0 * U[1, 4] per particle."""
    r = Rng(seed)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    counts = np.array([nx, ny, nz])
    sp = (hi - lo) / counts
    n = int(counts.prod())
    idx = np.indices((nx, ny, nz)).reshape(3, -1).T.astype(np.float64)
    jit = r.uniform(3 * n, -0.25, 0.25).reshape(n, 3)
    pos = lo + (idx + 0.5 + jit) * sp
    vel = r.normal(3 * n, sigma_v).reshape(n, 3)
    young = E0 * r.uniform(n, 1.0, 4.0)
    mass = RHO * float(np.prod(sp))
    desc = SolverDesc((0.0, 0.0, 0.0), (res, res, res), 1.0 / res, FRICTION, 0.95, N.QUADRATIC_BSPLINE, N.APIC, dt,
                      GRAVITY)
    return Scene("apic_slab", desc, snow_particles(pos, vel, mass, young=young), steps, seed,
                 f"{nx}x{ny}x{nz} jittered lattice, {res}^3, dt={dt}")


# ---------------------------------------------------------------- config 3
def primitive_union(rng: Rng, n_target: int, extent: float, center, kinds=None):
    """Seeded union of 1-4 spheres / boxes inside a cube of side ``extent``
    around ``center``; returns ~n_target points (uniform in the union)."""
    k = int(rng.integers(1, 1, 5)[0])
    prims = []
    for _ in range(k):
        kind = int(rng.integers(1, 0, 2)[0]) if kinds is None else kinds
        c = np.asarray(center) + rng.uniform(3, -0.2, 0.2) * extent
        s = extent * rng.uniform(1, 0.15, 0.3)[0]
        prims.append((kind, c, s))
    lo = np.min([p[1] - p[2] for p in prims], axis=0)
    hi = np.max([p[1] + p[2] for p in prims], axis=0)
    pts = []
    got = 0
    while got < n_target:
        m = max(2 * (n_target - got), 1024)
        cand = lo + rng.uniform(3 * m).reshape(m, 3) * (hi - lo)
        inside = np.zeros(m, bool)
        for kind, c, s in prims:
            if kind == 0:
                inside |= ((cand - c) ** 2).sum(1) <= s * s
            else:
                inside |= (np.abs(cand - c) <= s).all(1)
        sel = cand[inside]
        pts.append(sel)
        got += sel.shape[0]
    p = np.concatenate(pts)[:n_target]
    vol = float(np.prod(hi - lo)) * inside.mean() if inside.size else 1.0
    return p, max(vol, 1e-12)


def batch_scene(rng: Rng, index: int, n_lo: int = 2000, n_hi: int = 200000, steps: int = 200,
                transfer: int = N.APIC, interp: int = N.QUADRATIC_BSPLINE) -> Scene:
    """One scene of the uneven batch: N log-uniform in [n_lo, n_hi], grid
    32^3 / 64^3 / 128^3 chosen so the union spans ~40% of the domain at
    ~4-8 particles per cell."""
    r = rng.child(index)
    n = int(round(math.exp(r.uniform(1, math.log(n_lo), math.log(n_hi))[0])))
    res = 32 if n < 15000 else 64 if n < 60000 else 128
    pos, vol = primitive_union(r, n, 0.4, (0.5, 0.45, 0.5))
    vel = r.uniform(3, -1.0, 1.0)
    mass = RHO * vol / n
    desc = SolverDesc((0.0, 0.0, 0.0), (res, res, res), 1.0 / res, FRICTION, 0.95, interp, transfer, 1e-4, GRAVITY)
    return Scene(f"batch_{index}", desc, snow_particles(pos, vel, mass), steps, r.seed, f"N={n} res={res}")


def uneven_batch(seed: int = 3, count: int = 256, **kw):
    rng = Rng(seed)
    return [batch_scene(rng, i, **kw) for i in range(count)]


# ---------------------------------------------------------------- config 4
@dataclass
class Request:
    arrival: float
    scene: Scene


def request_stream(seed: int = 4, rate: float = 20.0, duration: float = 600.0, n_lo: int = 5000,
                   n_hi: int = 500000, steps_lo: int = 200, steps_hi: int = 2000, max_requests: int | None = None):
    """Poisson arrivals; each request a seeded scene (1-4 primitives,
    N log-uniform [n_lo, n_hi], grid 64^3-256^3 by size, PIC-FLIP or APIC
    50/50, steps uniform in [steps_lo, steps_hi])."""
    return list(iter_request_stream(seed, rate, duration, n_lo, n_hi, steps_lo, steps_hi, max_requests))


def iter_request_stream(seed: int = 4, rate: float = 20.0, duration: float = 600.0, n_lo: int = 5000,
                        n_hi: int = 500000, steps_lo: int = 200, steps_hi: int = 2000,
                        max_requests: int | None = None):
    """The requests of request_stream() one at a time, in arrival order."""
    rng = Rng(seed)
    t = 0.0
    i = 0
    while True:
        t += -math.log1p(-rng.uniform(1)[0]) / rate
        if t >= duration or (max_requests is not None and i >= max_requests):
            break
        r = rng.child(i)
        n = int(round(math.exp(r.uniform(1, math.log(n_lo), math.log(n_hi))[0])))
        res = 64 if n < 60000 else 128 if n < 250000 else 256
        transfer = N.APIC if r.uniform(1)[0] < 0.5 else N.PIC_FLIP
        steps = int(r.integers(1, steps_lo, steps_hi + 1)[0])
        pos, vol = primitive_union(r, n, 0.4, (0.5, 0.45, 0.5))
        vel = r.uniform(3, -1.0, 1.0)
        desc = SolverDesc((0.0, 0.0, 0.0), (res, res, res), 1.0 / res, FRICTION, 0.95, N.QUADRATIC_BSPLINE, transfer,
                          1e-4, GRAVITY)
        yield Request(t, Scene(f"req_{i}", desc, snow_particles(pos, vel, RHO * vol / n), steps, r.seed))
        i += 1


def workload(name: str, **kw) -> Scene:
    return {"snow_cube": snow_cube, "two_balls": two_balls, "apic_slab": apic_slab}[name](**kw)
