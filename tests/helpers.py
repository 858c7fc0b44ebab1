"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np


def rel_err(a, b) -> float:
    """Normwise relative error max|a-b| / max|b| (FP64)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def particle_errors(gpu_state, ref_state, fields=("position", "velocity", "Fe", "Fp")) -> dict:
    return {f: rel_err(getattr(gpu_state, f), getattr(ref_state, f)) for f in fields}


def grid_errors(g, r, fields=("mass", "velocity", "velocity_star", "force")) -> dict:
    return {f: rel_err(g[f], r[f]) for f in fields}


def report(tag: str, errs: dict):
    print(f"[{tag}] " + ", ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def particle_errors_apic(pa, pb, h) -> dict:
    """x, v, Fe, Fp normwise, and B relative to max|v| * h (B ~ v h, and B is
    ~0 for a near-uniform velocity field, where its own norm says nothing)."""
    e = particle_errors(pa, pb, ("position", "velocity", "Fe", "Fp"))
    e["B"] = float(np.abs(pa.B.astype(np.float64) - pb.B).max() / (np.abs(pb.velocity).max() * h + 1e-30))
    return e
