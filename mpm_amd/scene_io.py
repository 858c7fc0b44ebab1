"""Scene files at scale: the reference's YAML configuration
(src/mpm3d_parser.cpp:20-70) with text `.xyz` or binary `.xyzb` particle
groups, a loader that builds the solver inputs as the C++ builders do
(src/mpm_solver3d_builder.cu:10-79), and a generator for the benchmark
workloads.

Binary group layout (little-endian; read by the C++ builder too):

    0   char[8]  b"MPMXYZB1"
    8   int64    count n
    16  int32    fields: 1 velocity, 2 Young's modulus, 4 mass, 8 positions
                 in world units (else cell units, scaled by Grid.Stride)
    20  int32    0
    24  float32  position[n][3], then when flagged velocity[n][3], young[n],
                 mass[n]; per-particle fields replace the group's YAML values

Command line (writes a YAML + binary groups that `mpm_headless` runs):

    python -m mpm_amd.scene_io --workload apic_slab --out DIR
    python -m mpm_amd.scene_io --workload uneven_batch --count 256 --out DIR
    python -m mpm_amd.scene_io --workload request_stream --max-requests 100 --out DIR
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys

import numpy as np
import yaml

from . import native as N
from .solver import ParticleInput, SolverDesc

MAGIC = b"MPMXYZB1"
F_VEL, F_YOUNG, F_MASS, F_WORLD = 1, 2, 4, 8
INTERP_NAMES = {N.LINEAR: "Linear", N.QUADRATIC_BSPLINE: "QuadraticBSpline", N.CUBIC_BSPLINE: "CubicBSpline"}
MODEL_NAMES = {N.FIXED_COROTATED: "FixedCorotated", N.NEO_HOOKEAN: "NeoHookean"}


def write_xyzb(path: str, position, velocity=None, young=None, mass=None, world: bool = True):
    """Write one binary particle group."""
    pos = np.ascontiguousarray(position, "<f4").reshape(-1, 3)
    n = pos.shape[0]
    fields = F_WORLD if world else 0
    parts = [pos]
    for flag, arr, k in ((F_VEL, velocity, 3), (F_YOUNG, young, 1), (F_MASS, mass, 1)):
        if arr is not None:
            fields |= flag
            parts.append(np.ascontiguousarray(np.broadcast_to(np.asarray(arr, "<f4"), (n, k) if k > 1 else (n,))))
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([n], "<i8").tobytes())
        f.write(np.array([fields, 0], "<i4").tobytes())
        for a in parts:
            f.write(a.tobytes())


def read_xyzb(path: str) -> dict:
    """Read one binary particle group: position (as stored), world flag and
    the per-particle fields present."""
    raw = open(path, "rb").read()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not a binary particle group")
    n = int(np.frombuffer(raw, "<i8", 1, 8)[0])
    fields, zero = (int(v) for v in np.frombuffer(raw, "<i4", 2, 16))
    if n < 0 or zero != 0 or fields & ~15:
        raise ValueError(f"{path}: bad header")
    want = 24 + 4 * n * (3 + 3 * bool(fields & F_VEL) + bool(fields & F_YOUNG) + bool(fields & F_MASS))
    if len(raw) != want:
        raise ValueError(f"{path}: {len(raw)} bytes, expected {want}")
    off = 24
    out = {"world": bool(fields & F_WORLD)}

    def take(k):
        nonlocal off
        a = np.frombuffer(raw, "<f4", n * k, off).reshape((n, k) if k > 1 else (n,)).copy()
        off += 4 * n * k
        return a

    out["position"] = take(3)
    if fields & F_VEL:
        out["velocity"] = take(3)
    if fields & F_YOUNG:
        out["young"] = take(1)
    if fields & F_MASS:
        out["mass"] = take(1)
    return out


# A decimal number as a stream extraction of a float reads it (blanks before
# it; no inf, nan or hex): the C++ builder's scanner uses the same form.
_DECIMAL = re.compile(r"[ \t\r\v\f]*([+-]?(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?)")


def read_text_group(path: str):
    """(points float32 [n, 3], complete) of a text group: one "x y z" per
    line, anything after the third number ignored.  Reading stops at the
    first line without three numbers (a blank line included); the points
    before it are kept and complete is False (the reference's builder keeps
    them too, src/mpm_solver3d_builder.cu:10-30, :40)."""
    with open(path, "rb") as f:
        text = f.read().decode("latin-1")
    lines = text.split("\n")
    if text.endswith("\n") or not text:
        lines = lines[:-1]
    pts = []
    for line in lines:
        pos, xyz = 0, []
        for _ in range(3):
            m = _DECIMAL.match(line, pos)
            if not m:
                break
            v = np.float32(m.group(1))
            if np.isinf(v):
                break  # out of float range: the extraction fails
            xyz.append(v)
            pos = m.end()
        if len(xyz) < 3:
            return np.array(pts, np.float32).reshape(-1, 3), False
        pts.append(xyz)
    return np.array(pts, np.float32).reshape(-1, 3), True


def read_group(path: str, stride: float) -> dict:
    """A group as the builders read it: positions in world units (stride *
    coordinates for `.xyz` and cell-unit `.xyzb` files, in float32).  A text
    group that stops at a bad line holds the points before it."""
    if path.endswith(".xyzb"):
        g = read_xyzb(path)
    else:
        g = {"position": read_text_group(path)[0], "world": False}
    if not g.pop("world"):
        g["position"] = (np.float32(stride) * g["position"].astype(np.float32)).astype(np.float32)
    return g


def load_yaml(path: str, transfer: int | None = None, base_dir: str | None = None):
    """(SolverDesc, ParticleInput) from a configuration file.  Relative group
    paths resolve against base_dir, else the working directory (as the
    reference's builder).  transfer defaults to Transfer.Scheme, else
    PIC-FLIP.  An unreadable group raises (the C++ builder reports and skips
    it)."""
    with open(path) as f:
        cfg = yaml.safe_load(f)
    g = cfg["Grid"]
    stride = float(np.float32(g["Stride"]))
    sim = cfg["Simulate"]
    scheme = cfg["Transfer"].get("Scheme")
    if transfer is None:
        transfer = N.APIC if scheme == "APIC" else N.PIC_FLIP
    interp = {v: k for k, v in INTERP_NAMES.items()}[cfg["Interpolation"]["Type"]]
    desc = SolverDesc(tuple(float(v) for v in g["Origin"]), tuple(int(v) for v in g["Resolution"]), stride,
                      float(g["BoundaryFrictionCoefficient"]), float(cfg["Transfer"]["BlendCoefficient"]), interp,
                      transfer, float(sim["deltaTime"]), float(sim.get("Gravity", 9.80665)))
    groups = []
    for grp in cfg["ParticleGroups"]:
        fp = grp["File"]
        if base_dir is not None and not os.path.isabs(fp):
            fp = os.path.join(base_dir, fp)
        d = read_group(fp, stride)
        n = d["position"].shape[0]
        f32 = lambda v: np.full(n, np.float32(v), np.float32)
        groups.append(ParticleInput(
            d["position"], d.get("velocity", np.broadcast_to(np.float32(grp["Velocity"]), (n, 3))),
            d.get("mass", f32(grp["Mass"])), f32(grp["HardeningCoefficient"]), d.get("young", f32(grp["YoungModulus"])),
            f32(grp["PoissonRatio"]), f32(grp["CriticalCompression"]), f32(grp["CriticalStretch"]),
            np.full(n, {v: k for k, v in MODEL_NAMES.items()}[grp["ConstitutiveModel"]], np.int32)))
    return desc, ParticleInput.concat(groups)


def _uniform(a, name):
    a = np.asarray(a)
    if a.size and not np.all(a == a.flat[0]):
        raise ValueError(f"{name} varies per particle; the group format holds it per group")
    return a.flat[0].item() if a.size else 0


def _f(v) -> str:
    """A float32 value as the shortest decimal that parses back to it."""
    return str(np.float32(v))


def write_scene(desc: SolverDesc, particles: ParticleInput, out_dir: str, name: str, steps: int = 0) -> str:
    """Write `name`.yaml (the reference's layout: block mappings, flow
    sequences) and one binary group `name`.xyzb (world units, with
    per-particle velocity, Young's modulus and mass); returns the YAML path."""
    os.makedirs(out_dir, exist_ok=True)
    p = particles
    write_xyzb(os.path.join(out_dir, name + ".xyzb"), p.position, p.velocity, p.young, p.mass, world=True)
    vec = lambda v: "[" + ", ".join(_f(x) for x in v) + "]"
    lines = [
        "ParticleGroups:",
        f"  - File: {name}.xyzb",
        f"    Mass: {_f(p.mass[0]) if p.count else '0.0'}",
        "    Velocity: [0.0, 0.0, 0.0]",
        f"    YoungModulus: {_f(p.young[0]) if p.count else '0.0'}",
        f"    PoissonRatio: {_f(_uniform(p.poisson, 'PoissonRatio'))}",
        f"    HardeningCoefficient: {_f(_uniform(p.hardening, 'HardeningCoefficient'))}",
        f"    CriticalCompression: {_f(_uniform(p.critical_compression, 'CriticalCompression'))}",
        f"    CriticalStretch: {_f(_uniform(p.critical_stretch, 'CriticalStretch'))}",
        f"    ConstitutiveModel: {MODEL_NAMES[int(_uniform(p.model, 'ConstitutiveModel'))]}",
        "Grid:",
        f"  Origin: {vec(desc.origin)}",
        f"  Resolution: [{', '.join(str(int(v)) for v in desc.resolution)}]",
        f"  Stride: {_f(desc.stride)}",
        f"  BoundaryFrictionCoefficient: {_f(desc.friction)}",
        "Transfer:",
        f"  BlendCoefficient: {_f(desc.blend)}",
        f"  Scheme: {'APIC' if desc.transfer == N.APIC else 'PICFLIP'}",
        "Interpolation:",
        f"  Type: {INTERP_NAMES[desc.interp]}",
        "Simulate:",
        f"  deltaTime: {_f(desc.dt)}",
        f"  Gravity: {_f(desc.gravity)}",
        f"  Steps: {int(steps)}",
        "Render:",
        "  Offline: false",
        "  ParticleFragmentShader: ../shader/snow.frag",
    ]
    path = os.path.join(out_dir, name + ".yaml")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def main(argv=None) -> int:
    from . import scenes

    ap = argparse.ArgumentParser(description="write benchmark scenes as YAML + binary particle groups")
    ap.add_argument("--workload", required=True,
                    choices=["snow_cube", "two_balls", "apic_slab", "uneven_batch", "request_stream"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--count", type=int, default=256, help="uneven_batch: scenes")
    ap.add_argument("--max-requests", type=int, default=100, help="request_stream: requests written")
    ap.add_argument("--duration", type=float, default=600.0, help="request_stream: seconds of arrivals")
    a = ap.parse_args(argv)
    kw = {} if a.seed is None else {"seed": a.seed}
    if a.workload in ("snow_cube", "two_balls", "apic_slab"):
        sc = scenes.workload(a.workload, **kw)
        path = write_scene(sc.desc, sc.particles, a.out, sc.name, sc.steps)
        print(f"{path}: {sc.particles.count} particles ({sc.notes})")
        return 0
    if a.workload == "uneven_batch":
        index = []
        for sc in scenes.uneven_batch(count=a.count, **kw):
            index.append({"yaml": os.path.basename(write_scene(sc.desc, sc.particles, a.out, sc.name, sc.steps)),
                          "particles": sc.particles.count, "steps": sc.steps})
        with open(os.path.join(a.out, "batch.json"), "w") as f:
            json.dump(index, f, indent=1)
        print(f"{len(index)} scenes, {sum(e['particles'] for e in index)} particles -> {a.out}")
        return 0
    reqs = scenes.request_stream(duration=a.duration, max_requests=a.max_requests, **kw)
    os.makedirs(a.out, exist_ok=True)
    with open(os.path.join(a.out, "requests.jsonl"), "w") as f:
        for r in reqs:
            y = write_scene(r.scene.desc, r.scene.particles, a.out, r.scene.name, r.scene.steps)
            f.write(json.dumps({"arrival": r.arrival, "yaml": os.path.basename(y), "particles": r.scene.particles.count,
                                "steps": r.scene.steps}) + "\n")
    print(f"{len(reqs)} requests -> {a.out}/requests.jsonl")
    return 0


if __name__ == "__main__":
    sys.exit(main())
