"""The C++ drop-in end to end on the device: `mpm_headless` runs the
reference demo loop (parseYAML -> build{MPM,APICMPM}Solver3DFromYAML ->
simulateOneStep -> copyPositionsToHost, src/demo/mpm-apic/main.cu:47-209)
on the reference's own config/snow_cube.yaml and data/cube_sparse.xyz, and
its positions are compared with the FP32 oracle built from the same files
(group order, stride * xyz, per-group parameters:
src/mpm_solver3d_builder.cu:10-79)."""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd.solver import ParticleInput, SolverDesc
from oracle import pyoracle
from tests.helpers import rel_err, report

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG_DIR = os.path.join(ROOT, "tests", "golden", "reference_inputs", "config")
LIB_DIR = os.path.join(ROOT, "mpm_amd", "lib")


def _config(path):
    lib = C.CDLL(os.path.join(LIB_DIR, "libchains_mpm.so"))
    fn = lib.chains_parse_yaml_json
    fn.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    buf = C.create_string_buffer(1 << 16)
    assert fn(path.encode(), buf, len(buf)) == 0, buf.value
    return json.loads(buf.value.decode())


def _scene(cfg, transfer):
    groups = []
    for i, rel in enumerate(cfg["group_paths"]):
        xyz = np.loadtxt(os.path.join(CFG_DIR, rel), dtype=np.float64).reshape(-1, 3).astype(np.float32)
        n = xyz.shape[0]
        f32 = lambda v: np.full(n, np.float32(v), np.float32)
        groups.append(ParticleInput(
            np.float32(cfg["stride"]) * xyz, np.broadcast_to(np.float32(cfg["velocities"][i]), (n, 3)),
            f32(cfg["masses"][i]), f32(cfg["hardenings"][i]), f32(cfg["youngs"][i]), f32(cfg["poissons"][i]),
            f32(cfg["compressions"][i]), f32(cfg["stretches"][i]),
            np.full(n, cfg["constitutive_models"][i], np.int32)))
    particles = ParticleInput.concat(groups)
    desc = SolverDesc(tuple(cfg["origin"]), tuple(cfg["resolution"]), cfg["stride"], cfg["friction_coeff"],
                      cfg["blend_coeff"], cfg["interp_type"], transfer, cfg["delta_time"], cfg["gravity"])
    return desc, particles


@pytest.mark.parametrize("config,apic", [("snow_cube.yaml", False), ("snow_cube.yaml", True),
                                          ("snow_two_balls_sparse.yaml", False),
                                          ("snow_two_balls_sparse.yaml", True)])
def test_headless_reference_configs_match_oracle(config, apic, tmp_path):
    # the reference's two configurations whose particle files ship with it
    cfg_path = os.path.join(CFG_DIR, config)
    cfg = _config(cfg_path)
    steps = 2
    out = str(tmp_path / "pos.f32")
    cmd = [os.path.join(LIB_DIR, "mpm_headless"), cfg_path, "--steps", str(steps), "--out", out]
    if apic:
        cmd.append("--apic")
    print("running", " ".join(cmd), flush=True)
    # relative .xyz paths resolve against the working directory, as in the reference
    r = subprocess.run(cmd, cwd=CFG_DIR, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    pos = np.fromfile(out, dtype=np.float32).reshape(-1, 4)
    desc, particles = _scene(cfg, N.APIC if apic else N.PIC_FLIP)
    assert pos.shape[0] == particles.count == {"snow_cube.yaml": 11907, "snow_two_balls_sparse.yaml": 63008}[config]
    assert np.all(pos[:, 3] == 1.0)
    o = pyoracle.OracleSolver(desc, particles, 32)
    o.step(steps)
    ref = o.read_particles().position
    e = {"position": rel_err(pos[:, :3], ref), "moved": float(np.abs(ref - particles.position).max())}
    report(f"headless {config} {'APIC' if apic else 'PIC-FLIP'} {steps} steps", e)
    assert e["moved"] > 0.0
    assert e["position"] <= 1e-5


def test_headless_density_frames_match_oracle(tmp_path):
    # the offline output path: writeToOpenVDB every 2nd step into a directory
    cfg_path = os.path.join(CFG_DIR, "snow_cube.yaml")
    cfg = _config(cfg_path)
    out_dir = str(tmp_path / "frames")
    cmd = [os.path.join(LIB_DIR, "mpm_headless"), cfg_path, "--apic", "--steps", "4", "--every", "2",
           "--density", out_dir]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=CFG_DIR, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert sorted(os.listdir(out_dir)) == ["00002.mpmd", "00004.mpmd"]
    from mpm_amd.density import read_density_file
    desc, particles = _scene(cfg, N.APIC)
    o = pyoracle.OracleSolver(desc, particles, 32)
    o.step(4)
    st = o.read_particles()
    ref = pyoracle.density(desc, st.position, st.mass, 32)
    got = read_density_file(os.path.join(out_dir, "00004.mpmd"))
    assert got.resolution == tuple(cfg["resolution"])
    e = {"density": rel_err(got.dense(), ref)}
    report("headless density frame (APIC, 4 steps)", e)
    assert e["density"] <= 1e-5


@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_headless_runs_generated_binary_scene(tmp_path, transfer):
    # a generated scene written as YAML + binary group (per-particle velocity
    # and Young's modulus); Transfer.Scheme selects the solver class
    from mpm_amd import scenes
    from mpm_amd.scene_io import write_scene

    sc = scenes.snow_cube(seed=41, n=20000, res=48, transfer=transfer)
    rng = np.random.default_rng(4)
    sc.particles.velocity[:] = rng.normal(0, 0.5, (20000, 3)).astype(np.float32)
    sc.particles.young[:] = rng.uniform(1.4e5, 5.6e5, 20000).astype(np.float32)
    y = write_scene(sc.desc, sc.particles, str(tmp_path), "gen", steps=3)
    out = str(tmp_path / "pos.f32")
    cmd = [os.path.join(LIB_DIR, "mpm_headless"), y, "--out", out]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    pos = np.fromfile(out, dtype=np.float32).reshape(-1, 4)
    assert pos.shape[0] == sc.particles.count and np.all(pos[:, 3] == 1.0)
    # the FP32 oracle from the same inputs (what the binary group holds)
    o = pyoracle.OracleSolver(sc.desc, sc.particles, 32)
    o.step(3)
    ref = o.read_particles().position
    e = {"position": rel_err(pos[:, :3], ref), "moved": float(np.abs(ref - sc.particles.position).max())}
    report(f"headless generated scene {'APIC' if transfer else 'PIC-FLIP'} vs oracle", e)
    assert e["moved"] > 0.0 and e["position"] <= 1e-5


@pytest.mark.parametrize("config", ["snow_cube.yaml", "snow_two_balls_sparse.yaml"])
@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_reference_configs_full_state_vs_oracle(config, transfer):
    # the reference's configurations through the Python loader (the builder's
    # group order, stride * xyz, per-group parameters): velocities,
    # deformation gradients and (APIC) B after 3 steps, not only positions
    from mpm_amd.scene_io import load_yaml
    from mpm_amd.solver import Solver
    from tests.helpers import particle_errors, particle_errors_apic

    desc, particles = load_yaml(os.path.join(CFG_DIR, config), transfer=transfer, base_dir=CFG_DIR)
    gpu = Solver(desc, particles)
    ref = pyoracle.OracleSolver(desc, particles, 32)
    gpu.step(3)
    ref.step(3)
    g, r = gpu.read_particles(), ref.read_particles()
    e = particle_errors_apic(g, r, desc.stride) if transfer == N.APIC else particle_errors(g, r)
    e["volume0"] = rel_err(g.volume0, r.volume0)
    report(f"{config} {'APIC' if transfer else 'PIC-FLIP'} 3 steps, full state", e)
    assert all(v <= 1e-5 for v in e.values()), e
