"""ctypes wrapper of the CPU ORACLE (liboracle_mpm.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the solver package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from mpm_amd import native as N
from mpm_amd.solver import ParticleInput, ParticleState, SolverDesc

HERE = os.path.dirname(os.path.abspath(__file__))
# MPM_ORACLE_LIB names another build of the same sources (the mutation test
# of the oracle's pins loads deliberately broken copies this way).
LIB = os.environ.get("MPM_ORACLE_LIB") or os.path.join(HERE, "liboracle_mpm.so")

_lib = None
_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_create.restype = C.c_void_p
    lib.orc_create.argtypes = [C.POINTER(N.MpmDesc), C.POINTER(N.MpmParticleInput), C.c_int]
    lib.orc_create_threads.restype = C.c_void_p
    lib.orc_create_threads.argtypes = [C.POINTER(N.MpmDesc), C.POINTER(N.MpmParticleInput), C.c_int, C.c_int]
    lib.orc_create_state.restype = C.c_void_p
    lib.orc_create_state.argtypes = [C.POINTER(N.MpmDesc), C.POINTER(N.MpmParticleState), C.c_int, C.c_int]
    lib.orc_destroy.argtypes = [C.c_void_p]
    lib.orc_step.argtypes = [C.c_void_p, C.c_int]
    lib.orc_substep.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float]
    lib.orc_read_particles.argtypes = [C.c_void_p, C.POINTER(N.MpmParticleState)]
    lib.orc_read_particles64.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    lib.orc_read_grid.argtypes = [C.c_void_p, C.POINTER(N.MpmGridHost)]
    lib.orc_set_particle_collision.argtypes = [C.c_void_p, C.c_int]
    lib.orc_time_steps.restype = C.c_double
    lib.orc_time_steps.argtypes = [C.c_void_p, C.c_int]
    lib.orc_svd3.argtypes = [_dp, C.c_int64, _dp, _dp, _dp, C.c_int]
    lib.orc_weights.argtypes = [C.c_int, _dp, C.c_int64, _dp, _dp, C.c_int]
    lib.orc_stress.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp, C.c_int]
    lib.orc_update_F.argtypes = [_dp, _dp, _dp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int]
    lib.orc_collide.argtypes = [C.POINTER(C.c_float)] * 4 + [C.c_float, C.POINTER(C.c_float)]
    lib.orc_density.argtypes = [C.POINTER(N.MpmDesc), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, _dp,
                                C.POINTER(C.c_uint8), C.c_int]
    lib.orc_time_density.restype = C.c_double
    lib.orc_time_density.argtypes = [C.POINTER(N.MpmDesc), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64]
    lib.orc_keys.argtypes = [C.POINTER(N.MpmDesc), C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_uint32)]
    lib.orc_stable_argsort.argtypes = [C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_uint32)]
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


SUBSTEPS = {"reset": 0, "p2g": 1, "gravity": 2, "update": 3, "collision": 4, "g2p": 5, "pcollision": 6,
            "advect": 7, "initialize": 8}


class OracleSolver:
    """Serial CPU restatement of the reference step (precision 32 or 64).

    threads > 1 (inputs only; the all-cores timing) runs the loops under
    OpenMP with atomic grid sums, whose order varies from run to run."""

    def __init__(self, desc: SolverDesc, particles, precision: int = 32, run_initialize: bool = True,
                 threads: int = 1):
        self._lib = load()
        self.desc = desc
        self._cdesc = desc.to_c()
        if threads != 1 and isinstance(particles, ParticleState):
            raise ValueError("threaded oracle takes solver inputs only")
        if isinstance(particles, ParticleState):
            st = particles.copy()
            c = st.to_c()
            h = self._lib.orc_create_state(C.byref(self._cdesc), C.byref(c), 1 if run_initialize else 0, precision)
        else:
            c = particles.to_c()
            h = self._lib.orc_create_threads(C.byref(self._cdesc), C.byref(c), precision, int(threads))
        if not h:
            raise ValueError("oracle rejected the scene")
        self._h = C.c_void_p(h)
        self.count = particles.count
        self.precision = precision

    def close(self):
        if self._h and self._h.value:
            self._lib.orc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        self._lib.orc_step(self._h, n)

    def time_steps(self, n: int) -> float:
        return self._lib.orc_time_steps(self._h, n)

    def substep(self, name: str, dt: float = 0.0, blend: float = 0.0):
        self._lib.orc_substep(self._h, SUBSTEPS[name], dt, blend)

    def set_particle_collision(self, on: bool):
        self._lib.orc_set_particle_collision(self._h, 1 if on else 0)

    def read_particles(self) -> ParticleState:
        st = ParticleState.empty(self.count)
        c = st.to_c()
        self._lib.orc_read_particles(self._h, C.byref(c))
        return st

    def read_particles64(self) -> dict:
        n = self.count
        out = {k: np.zeros(s, np.float64) for k, s in
               (("position", (n, 3)), ("velocity", (n, 3)), ("Fe", (n, 3, 3)), ("Fp", (n, 3, 3)), ("B", (n, 3, 3)))}
        self._lib.orc_read_particles64(self._h, _d(out["position"]), _d(out["velocity"]), _d(out["Fe"]),
                                       _d(out["Fp"]), _d(out["B"]))
        return out

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
        self._lib.orc_read_grid(self._h, C.byref(c))
        return out


def svd3(A: np.ndarray, precision: int = 64):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(-1, 3, 3)
    n = A.shape[0]
    U = np.zeros((n, 3, 3))
    V = np.zeros((n, 3, 3))
    s = np.zeros((n, 3))
    lib.orc_svd3(_d(A), n, _d(U), _d(s), _d(V), precision)
    return U, s, V


def weights(interp: int, x, precision: int = 64):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    Nv = np.zeros_like(x)
    dN = np.zeros_like(x)
    lib.orc_weights(interp, _d(x), x.size, _d(Nv), _d(dN), precision)
    return Nv, dN


def stress(Fe, Fp, mu0, lambda0, xi, model=N.FIXED_COROTATED, precision: int = 64) -> np.ndarray:
    lib = load()
    Fe = np.ascontiguousarray(Fe, dtype=np.float64).reshape(3, 3)
    Fp = np.ascontiguousarray(Fp, dtype=np.float64).reshape(3, 3)
    out = np.zeros((3, 3))
    lib.orc_stress(_d(Fe), _d(Fp), mu0, lambda0, xi, model, _d(out), precision)
    return out


def update_F(Fe, Fp, grad, dt, crit_c, crit_s, precision: int = 64):
    """updateDeformationGradient of a batch: returns the new (Fe, Fp)."""
    lib = load()
    Fe = np.array(Fe, dtype=np.float64).reshape(-1, 3, 3)
    Fp = np.array(Fp, dtype=np.float64).reshape(-1, 3, 3)
    g = np.ascontiguousarray(grad, dtype=np.float64).reshape(-1, 3, 3)
    lib.orc_update_F(_d(Fe), _d(Fp), _d(g), Fe.shape[0], dt, crit_c, crit_s, precision)
    return Fe, Fp


def collide(pos, vel, g000, g111, mu) -> np.ndarray:
    lib = load()
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    p, v, a, b = f(pos), f(vel), f(g000), f(g111)
    out = np.zeros(3, np.float32)
    lib.orc_collide(N.fptr(p), N.fptr(v), N.fptr(a), N.fptr(b), mu, N.fptr(out))
    return out


def keys(desc: SolverDesc, positions) -> np.ndarray:
    lib = load()
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(p.shape[0], np.uint32)
    d = desc.to_c()
    lib.orc_keys(C.byref(d), N.fptr(p), p.shape[0], N.uptr(out))
    return out


def stable_argsort(k) -> np.ndarray:
    lib = load()
    k = np.ascontiguousarray(k, dtype=np.uint32)
    out = np.zeros(k.size, np.uint32)
    lib.orc_stable_argsort(N.uptr(k), k.size, N.uptr(out))
    return out


def density(desc: SolverDesc, positions, mass, precision: int = 32, with_active: bool = False):
    """writeToOpenVDB's density raster of the given particles: flat array in
    the reference node order (and the written-node mask if asked)."""
    lib = load()
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (p.shape[0],)))
    d = desc.to_c()
    out = np.zeros(desc.nodes, np.float64)
    act = np.zeros(desc.nodes, np.uint8) if with_active else None
    lib.orc_density(C.byref(d), N.fptr(p), N.fptr(m), p.shape[0], _d(out),
                    act.ctypes.data_as(C.POINTER(C.c_uint8)) if with_active else None, precision)
    return (out, act.astype(bool)) if with_active else out


def time_density(desc: SolverDesc, positions, mass) -> float:
    """Seconds for one serial FP32 raster of the given particles."""
    lib = load()
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (p.shape[0],)))
    d = desc.to_c()
    return float(lib.orc_time_density(C.byref(d), N.fptr(p), N.fptr(m), p.shape[0]))
