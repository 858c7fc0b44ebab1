"""Many scenes advanced together: Python side of the ``mpm_batch_*`` C-ABI.

A :class:`Batch` groups :class:`~mpm_amd.solver.Solver` handles of one
interpolation and transfer scheme.  ``step(n)`` advances every member by n
steps with one batched clear, one radix sort over all scenes, one P2G and
one G2P launch per step (``step_graph`` replays the step as a hipGraph).
Members stay ordinary solver handles: read them, step them alone, remove
them.  The reference has one solver object per scene
(src/mpm_solver3d.cu:77-108) and no batching; this module is the MI355X
addition that the scene-batching and serving workloads run on.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as N
from .solver import Solver


@dataclass
class Tune:
    item_max: int = 256
    g2p_threads: int = 128
    workgroups: int = 0
    bucket: int = 0
    step_ms: float = 0.0
    fuse_transfer: int = 0  # 1: G2P fused with the next P2G (APIC quadratic batches)

    def to_c(self) -> N.MpmTune:
        return N.MpmTune(self.item_max, self.g2p_threads, self.workgroups, self.bucket, self.step_ms,
                         self.fuse_transfer)

    @staticmethod
    def from_c(t: N.MpmTune) -> "Tune":
        return Tune(t.item_max, t.g2p_threads, t.workgroups, t.bucket, t.step_ms, t.fuse_transfer)


class Batch:
    """A set of scenes stepped by shared launches (see module docstring)."""

    def __init__(self, interp: int, transfer: int):
        self._lib = N.load()
        self.interp = int(interp)
        self.transfer = int(transfer)
        h = C.c_void_p()
        N.check(self._lib.mpm_batch_create(self.interp, self.transfer, C.byref(h)))
        self._h = h
        self.members: list[Solver] = []

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.mpm_batch_destroy(self._h)
            self._h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.members)

    def add(self, s: Solver):
        N.check(self._lib.mpm_batch_add(self._h, s._h))
        self.members.append(s)

    def remove(self, s: Solver):
        N.check(self._lib.mpm_batch_remove(self._h, s._h))
        self.members = [m for m in self.members if m is not s]

    def step(self, n: int = 1):
        N.check(self._lib.mpm_batch_step(self._h, int(n)))

    def step_graph(self, n: int = 1):
        N.check(self._lib.mpm_batch_step_graph(self._h, int(n)))

    def synchronize(self):
        N.check(self._lib.mpm_batch_synchronize(self._h))

    def set_reorder_interval(self, steps: int):
        """Permute every member's store into cell order every `steps` batch
        steps (0 = never; default 16)."""
        N.check(self._lib.mpm_batch_set_reorder_interval(self._h, int(steps)))

    def set_incremental_sort(self, enable: bool):
        """Update the last step's cell order instead of sorting in full (on
        by default; bit-identical to the full sort)."""
        N.check(self._lib.mpm_batch_set_incremental_sort(self._h, int(bool(enable))))

    def read_sort(self):
        """(keys, global slots) of the last step's sorted order."""
        n = int(self.info()["particles"])
        keys = np.empty(n, np.uint32)
        slots = np.empty(n, np.uint32)
        N.check(self._lib.mpm_batch_read_sort(self._h, N.uptr(keys), N.uptr(slots), n))
        return keys, slots

    def info(self) -> dict:
        s = N.MpmBatchStats()
        N.check(self._lib.mpm_batch_info(self._h, C.byref(s)))
        d = {f[0]: getattr(s, f[0]) for f in N.MpmBatchStats._fields_ if f[0] != "tune"}
        d["tune"] = Tune.from_c(s.tune)
        return d

    def set_tune(self, t: Tune | None):
        if t is None:
            N.check(self._lib.mpm_batch_set_tune(self._h, None))
        else:
            c = t.to_c()
            N.check(self._lib.mpm_batch_set_tune(self._h, C.byref(c)))

    def autotune(self, trial_steps: int = 3) -> Tune:
        out = N.MpmTune()
        N.check(self._lib.mpm_batch_autotune(self._h, int(trial_steps), C.byref(out)))
        return Tune.from_c(out)

    def health(self) -> list[dict]:
        """Per member (membership order): max_speed, cfl, nonfinite and
        escaped particle counts of the current state (one device pass)."""
        n = len(self.members)
        arr = (N.MpmSceneHealth * max(n, 1))()
        N.check(self._lib.mpm_batch_health(self._h, arr, n))
        return [{f[0]: getattr(arr[i], f[0]) for f in N.MpmSceneHealth._fields_} for i in range(n)]

    def device_bytes(self) -> int:
        return int(self.info()["device_bytes"])

    def timing(self, enable: bool = True):
        N.check(self._lib.mpm_batch_timing(self._h, 1 if enable else 0))

    def read_times(self) -> np.ndarray:
        cnt = C.c_int()
        N.check(self._lib.mpm_batch_read_times(self._h, None, 0, C.byref(cnt)))  # count only
        return np.zeros(0, np.float32) if cnt.value == 0 else self._read_times_sized(cnt.value)

    def _read_times_sized(self, k: int) -> np.ndarray:
        out = np.zeros(k, np.float32)
        cnt = C.c_int()
        N.check(self._lib.mpm_batch_read_times(self._h, N.fptr(out), k, C.byref(cnt)))
        return out[: min(k, cnt.value)]


def tune_lookup(interp: int, transfer: int, particles: int) -> Tune | None:
    lib = N.load()
    out = N.MpmTune()
    rc = lib.mpm_tune_lookup(int(interp), int(transfer), int(particles), C.byref(out))
    if rc == N.MPM_ERR_NOT_FOUND:
        return None
    N.check(rc)
    return Tune.from_c(out)
