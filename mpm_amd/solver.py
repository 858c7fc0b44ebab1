"""Python mirror of the reference solver classes over the C-ABI.

``MPMSolver3D`` / ``APICMPMSolver3D`` keep the reference's method names
(reference include/mpm_solver3d.h:21-96, include/apic_mpm_solver3d.h:20-92);
each call forwards to one ``mpm_*`` entry point of ``libmpm_hip.so``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields

import numpy as np

from . import native as N
from .density import DensityBlocks


@dataclass
class SolverDesc:
    """Constructor arguments of the reference solvers (plus gravity)."""

    origin: tuple = (0.0, 0.0, 0.0)
    resolution: tuple = (64, 64, 64)
    stride: float = 1.0 / 64
    friction: float = 0.9
    blend: float = 0.95
    interp: int = N.QUADRATIC_BSPLINE
    transfer: int = N.PIC_FLIP
    dt: float = 1e-4
    gravity: float = 9.80665
    particle_collision: bool = False

    def to_c(self) -> N.MpmDesc:
        d = N.MpmDesc()
        for i in range(3):
            d.origin[i] = float(self.origin[i])
            d.resolution[i] = int(self.resolution[i])
        d.stride = float(self.stride)
        d.friction = float(self.friction)
        d.blend = float(self.blend)
        d.interp = int(self.interp)
        d.transfer = int(self.transfer)
        d.dt = float(self.dt)
        d.gravity = float(self.gravity)
        d.particle_collision = 1 if self.particle_collision else 0
        return d

    @property
    def nodes(self) -> int:
        return int(np.prod(self.resolution))


def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


@dataclass
class ParticleInput:
    """Per-particle MaterialPoint3D constructor arguments
    (reference include/material_point3d.h:27-37)."""

    position: np.ndarray
    velocity: np.ndarray
    mass: np.ndarray
    hardening: np.ndarray
    young: np.ndarray
    poisson: np.ndarray
    critical_compression: np.ndarray
    critical_stretch: np.ndarray
    model: np.ndarray

    def __post_init__(self):
        n = np.asarray(self.position).reshape(-1, 3).shape[0]
        self.position = _f32(self.position, (n, 3))
        self.velocity = _f32(self.velocity, (n, 3))
        for name in ("mass", "hardening", "young", "poisson", "critical_compression", "critical_stretch"):
            v = getattr(self, name)
            if np.ndim(v) == 0:
                v = np.full(n, v)
            setattr(self, name, _f32(v, (n,)))
        m = self.model
        if np.ndim(m) == 0:
            m = np.full(n, m)
        self.model = np.ascontiguousarray(np.asarray(m, dtype=np.int32).reshape(n))

    @property
    def count(self) -> int:
        return self.position.shape[0]

    def to_c(self) -> N.MpmParticleInput:
        c = N.MpmParticleInput()
        c.count = self.count
        c.position = N.fptr(self.position)
        c.velocity = N.fptr(self.velocity)
        c.mass = N.fptr(self.mass)
        c.hardening = N.fptr(self.hardening)
        c.young = N.fptr(self.young)
        c.poisson = N.fptr(self.poisson)
        c.critical_compression = N.fptr(self.critical_compression)
        c.critical_stretch = N.fptr(self.critical_stretch)
        c.model = N.iptr(self.model)
        return c

    def subset(self, idx) -> "ParticleInput":
        return ParticleInput(**{f.name: getattr(self, f.name)[idx] for f in fields(self)})

    @staticmethod
    def concat(parts) -> "ParticleInput":
        return ParticleInput(**{f.name: np.concatenate([getattr(p, f.name) for p in parts]) for f in fields(ParticleInput)})


@dataclass
class ParticleState:
    """Full MaterialPoint3D (+ APIC B) state, original particle order."""

    volume0: np.ndarray
    mass: np.ndarray
    position: np.ndarray
    velocity: np.ndarray
    Fe: np.ndarray
    Fp: np.ndarray
    mu0: np.ndarray
    lambda0: np.ndarray
    hardening: np.ndarray
    critical_compression: np.ndarray
    critical_stretch: np.ndarray
    model: np.ndarray
    B: np.ndarray = field(default=None)

    @staticmethod
    def empty(n: int) -> "ParticleState":
        z = lambda *s: np.zeros(s, dtype=np.float32)
        return ParticleState(z(n), z(n), z(n, 3), z(n, 3), z(n, 3, 3), z(n, 3, 3), z(n), z(n), z(n), z(n), z(n),
                             np.zeros(n, dtype=np.int32), z(n, 3, 3))

    @property
    def count(self) -> int:
        return self.position.shape[0]

    def to_c(self) -> N.MpmParticleState:
        c = N.MpmParticleState()
        c.count = self.count
        for name in ("volume0", "mass", "position", "velocity", "Fe", "Fp", "mu0", "lambda0", "hardening",
                     "critical_compression", "critical_stretch"):
            a = getattr(self, name)
            setattr(self, name, np.ascontiguousarray(a, dtype=np.float32))
            setattr(c, name, N.fptr(getattr(self, name)))
        self.model = np.ascontiguousarray(self.model, dtype=np.int32)
        c.model = N.iptr(self.model)
        if self.B is not None:
            self.B = np.ascontiguousarray(self.B, dtype=np.float32)
            c.B = N.fptr(self.B)
        return c

    def copy(self) -> "ParticleState":
        return ParticleState(**{f.name: (None if getattr(self, f.name) is None else getattr(self, f.name).copy())
                                for f in fields(self)})


class Solver:
    """One device solver handle (one scene)."""

    def __init__(self, desc: SolverDesc, particles, run_initialize: bool = True):
        self._lib = N.load()
        self.desc = desc
        self._cdesc = desc.to_c()
        h = C.c_void_p()
        if isinstance(particles, ParticleState):
            cst = particles.to_c()
            N.check(self._lib.mpm_create_state(C.byref(self._cdesc), C.byref(cst), 1 if run_initialize else 0,
                                               C.byref(h)))
        else:
            cin = particles.to_c()
            N.check(self._lib.mpm_create(C.byref(self._cdesc), C.byref(cin), C.byref(h)))
        self._h = h
        self.count = particles.count

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mpm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference methods (include/mpm_solver3d.h:37-63)
    def simulateOneStep(self, n: int = 1):
        N.check(self._lib.mpm_step(self._h, int(n)))

    step = simulateOneStep

    def step_graph(self, n: int = 1):
        N.check(self._lib.mpm_step_graph(self._h, int(n)))

    def synchronize(self):
        N.check(self._lib.mpm_synchronize(self._h))

    def switchIfEnableParticlesCollision(self):
        N.check(self._lib.mpm_switch_particles_collision(self._h))

    def set_particles_collision(self, on: bool):
        N.check(self._lib.mpm_set_particles_collision(self._h, 1 if on else 0))

    def set_grid_record(self, on: bool):
        """APIC: keep momentum and force apart in the fused step so that
        read_grid works after it (default: folded, faster)."""
        N.check(self._lib.mpm_set_grid_record(self._h, 1 if on else 0))

    def set_reorder_interval(self, steps: int):
        """Permute the particle store into cell order every `steps` steps
        (0 = never; default 100)."""
        N.check(self._lib.mpm_set_reorder_interval(self._h, int(steps)))

    def set_transfer_fusion(self, on: bool):
        """APIC, quadratic stencil, stepped alone: G2P of one step and P2G
        of the next in one kernel (default on)."""
        N.check(self._lib.mpm_set_transfer_fusion(self._h, 1 if on else 0))

    def initialize(self):
        N.check(self._lib.mpm_initialize(self._h))

    def resetGrid(self):
        N.check(self._lib.mpm_reset_grid(self._h))

    def particlesToGrid(self):
        N.check(self._lib.mpm_particles_to_grid(self._h))

    def computeExternalForces(self):
        N.check(self._lib.mpm_compute_external_forces(self._h))

    def computeGravityForces(self):
        N.check(self._lib.mpm_compute_gravity_forces(self._h))

    def updateGridVelocities(self, dt: float):
        N.check(self._lib.mpm_update_grid_velocities(self._h, float(dt)))

    def solveLinearSystem(self, dt: float):
        N.check(self._lib.mpm_solve_linear_system(self._h, float(dt)))

    def gridCollision(self, dt: float):
        N.check(self._lib.mpm_grid_collision(self._h, float(dt)))

    def gridToParticles(self, dt: float, blend: float | None = None):
        b = self.desc.blend if blend is None else blend
        N.check(self._lib.mpm_grid_to_particles(self._h, float(dt), float(b)))

    def particlesCollision(self, dt: float):
        N.check(self._lib.mpm_particles_collision(self._h, float(dt)))

    def advectParticles(self, dt: float):
        N.check(self._lib.mpm_advect_particles(self._h, float(dt)))

    def sort(self):
        N.check(self._lib.mpm_sort(self._h))

    # -- state I/O
    def read_particles(self, noalloc: bool = False) -> ParticleState:
        """State in the caller's particle order.  noalloc: stage through the
        library's static device array (no hipMalloc; at most 262,144
        particles)."""
        st = ParticleState.empty(self.count)
        if self.desc.transfer != N.APIC:
            st.B = None
        c = st.to_c()
        fn = self._lib.mpm_read_particles_noalloc if noalloc else self._lib.mpm_read_particles
        N.check(fn(self._h, C.byref(c)))
        return st

    def write_particles(self, st: ParticleState):
        c = st.to_c()
        N.check(self._lib.mpm_write_particles(self._h, C.byref(c)))

    def read_grid(self) -> dict:
        g = self.desc.nodes
        out = {"mass": np.zeros(g, np.float32), "velocity": np.zeros((g, 3), np.float32),
               "velocity_star": np.zeros((g, 3), np.float32), "force": np.zeros((g, 3), np.float32)}
        c = N.MpmGridHost()
        c.nodes = g
        c.mass = N.fptr(out["mass"])
        c.velocity = N.fptr(out["velocity"])
        c.velocity_star = N.fptr(out["velocity_star"])
        c.force = N.fptr(out["force"])
        N.check(self._lib.mpm_read_grid(self._h, C.byref(c)))
        return out

    def read_positions4(self) -> np.ndarray:
        out = np.zeros((self.count, 4), np.float32)
        N.check(self._lib.mpm_read_positions4(self._h, N.fptr(out)))
        return out

    def pack_positions4(self, device_ptr: int):
        """Pack (x, y, z, 1) into a device buffer of 4*count floats
        (asynchronous on the solver's stream)."""
        N.check(self._lib.mpm_pack_positions4(self._h, C.c_void_p(int(device_ptr))))

    def positions4_async(self, out: np.ndarray):
        """Start a frame export into `out` ((count, 4) float32, C-contiguous);
        it is complete after positions4_wait()."""
        if out.dtype != np.float32 or out.shape != (self.count, 4) or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous (count, 4) float32 array")
        self._pos4_out = out  # keep the buffer alive until the copy is waited on
        N.check(self._lib.mpm_positions4_async(self._h, N.fptr(out)))

    def positions4_wait(self):
        N.check(self._lib.mpm_positions4_wait(self._h))
        self._pos4_out = None

    # ---- density raster (replacement of writeToOpenVDB, src/mpm_solver3d.cu:510-572)
    def density_compute(self):
        """Rasterise the current positions on the device (asynchronous)."""
        N.check(self._lib.mpm_density_compute(self._h))

    def density(self) -> np.ndarray:
        """Dense density m w / h^3 per node, flat in the reference node order
        x + rx*(y + ry*z)."""
        out = np.zeros(self.desc.nodes, np.float32)
        N.check(self._lib.mpm_read_density(self._h, N.fptr(out)))
        return out

    def density_blocks(self) -> DensityBlocks:
        """The 4x4x4 node blocks the raster touched."""
        cnt = C.c_int64(0)
        N.check(self._lib.mpm_read_density_blocks(self._h, None, None, 0, C.byref(cnt)))
        n = cnt.value
        first = np.zeros((n, 3), np.int32)
        vals = np.zeros((n, 4, 4, 4), np.float32)
        if n:
            N.check(self._lib.mpm_read_density_blocks(
                self._h, first.ctypes.data_as(C.POINTER(C.c_int32)), N.fptr(vals), n, C.byref(cnt)))
        d = self.desc
        return DensityBlocks(tuple(d.resolution), tuple(d.origin), float(d.stride), first, vals)

    def write_density(self, path: str):
        """Write the density block file (layout in mpm_amd.density)."""
        N.check(self._lib.mpm_write_density(self._h, str(path).encode()))

    def read_sort(self):
        n = self.count
        keys = np.zeros(n, np.uint32)
        ids = np.zeros(n, np.uint32)
        seq = np.zeros(n, np.uint32)
        N.check(self._lib.mpm_read_sort(self._h, N.uptr(keys), N.uptr(ids), N.uptr(seq), n))
        return keys, ids, seq

    def host_keys(self, positions: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.uint32)
        N.check(self._lib.mpm_host_keys(self._h, N.fptr(p), p.shape[0], N.uptr(out)))
        return out

    def stats(self) -> dict:
        s = N.MpmStats()
        N.check(self._lib.mpm_get_stats(self._h, C.byref(s)))
        return {f[0]: getattr(s, f[0]) for f in N.MpmStats._fields_}

    def profile(self, enable=True):
        """enable: False/0 off, True/1 every kernel group, 2 only the item
        kernels (p2g, g2p, g2p2g) and the store reorder (fewer events in
        the stream)."""
        N.check(self._lib.mpm_profile_enable(self._h, int(enable)))

    def profile_reset(self):
        N.check(self._lib.mpm_profile_reset(self._h))

    def profile_read(self) -> dict:
        arr = (N.MpmKernelTime * 64)()
        cnt = C.c_int()
        N.check(self._lib.mpm_profile_read(self._h, arr, 64, C.byref(cnt)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(min(cnt.value, 64))}


class MPMSolver3D(Solver):
    """PIC-FLIP solver (reference include/mpm_solver3d.h:25-34)."""

    def __init__(self, particles, origin, resolution, stride, friction, blend, interp, dt, gravity=9.80665,
                 run_initialize=True):
        d = SolverDesc(tuple(origin), tuple(resolution), stride, friction, blend, interp, N.PIC_FLIP, dt, gravity)
        super().__init__(d, particles, run_initialize)

    def gridToParticles(self, dt: float, blend: float | None = None):
        super().gridToParticles(dt, blend)


class APICMPMSolver3D(Solver):
    """APIC solver (reference include/apic_mpm_solver3d.h:24-32)."""

    def __init__(self, particles, origin, resolution, stride, friction, interp, dt, gravity=9.80665,
                 run_initialize=True):
        d = SolverDesc(tuple(origin), tuple(resolution), stride, friction, 0.0, interp, N.APIC, dt, gravity)
        super().__init__(d, particles, run_initialize)


def debug_svd3(A: np.ndarray):
    """The device SVD on n row-major 3x3 matrices -> (U, s, V)."""
    lib = N.load()
    A = np.ascontiguousarray(A, dtype=np.float32).reshape(-1, 3, 3)
    n = A.shape[0]
    U = np.zeros((n, 3, 3), np.float32)
    V = np.zeros((n, 3, 3), np.float32)
    s = np.zeros((n, 3), np.float32)
    N.check(lib.mpm_debug_svd3(N.fptr(A), n, N.fptr(U), N.fptr(s), N.fptr(V)))
    return U, s, V


def cache_info() -> dict:
    """Device block cache of the current device: bytes cached and held, and
    allocations served from the cache (hits) or from hipMalloc (misses)."""
    lib = N.load()
    out = N.MpmCacheInfo()
    N.check(lib.mpm_get_cache_info(C.byref(out)))
    return {"cached_bytes": out.cached_bytes, "live_bytes": out.live_bytes, "hits": out.hits, "misses": out.misses}


def set_cache_limit(nbytes: int):
    """Bytes of released device blocks kept per device for reuse."""
    N.check(N.load().mpm_set_cache_limit(int(nbytes)))


def release_cache():
    """Return every cached device block to the driver (waits for the device)."""
    N.check(N.load().mpm_release_cache())
