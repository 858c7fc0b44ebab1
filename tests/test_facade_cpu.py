"""CPU checks of the drop-in boundary: the C++ YAML reader against PyYAML on
the reference's configuration files, the C-ABI export list against
include/mpm_hip.h, and argument validation paths that never reach a GPU."""
import ctypes as C
import glob
import json
import os
import re

import numpy as np
import pytest
import yaml

from mpm_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG_DIR = os.path.join(ROOT, "tests", "golden", "reference_inputs", "config")
FACADE_LIB = os.path.join(ROOT, "mpm_amd", "lib", "libchains_mpm.so")


def _parse_cpp(path):
    lib = C.CDLL(FACADE_LIB)
    fn = lib.chains_parse_yaml_json
    fn.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    buf = C.create_string_buffer(1 << 16)
    rc = fn(path.encode(), buf, len(buf))
    return rc, buf.value.decode()


INTERP = {"Linear": 0, "QuadraticBSpline": 1, "CubicBSpline": 2}
MODEL = {"FixedCorotated": 0, "NeoHookean": 1}


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(CFG_DIR, "*.yaml"))))
def test_parser_matches_pyyaml(path):
    rc, text = _parse_cpp(path)
    assert rc == 0, text
    got = json.loads(text)
    # 9 significant digits round-trip float32 exactly: compare as float32
    for k in ("masses", "youngs", "poissons", "hardenings", "compressions", "stretches", "origin"):
        got[k] = [float(np.float32(v)) for v in got[k]]
    got["velocities"] = [[float(np.float32(v)) for v in vel] for vel in got["velocities"]]
    for k in ("stride", "friction_coeff", "blend_coeff", "delta_time"):
        got[k] = float(np.float32(got[k]))
    with open(path) as f:
        y = yaml.safe_load(f)
    g = y["ParticleGroups"]
    f32 = lambda v: float(np.float32(v))
    assert got["group_paths"] == [p["File"] for p in g]
    assert got["masses"] == [f32(p["Mass"]) for p in g]
    assert got["velocities"] == [[f32(v) for v in p["Velocity"]] for p in g]
    assert got["youngs"] == [f32(p["YoungModulus"]) for p in g]
    assert got["poissons"] == [f32(p["PoissonRatio"]) for p in g]
    assert got["hardenings"] == [f32(p["HardeningCoefficient"]) for p in g]
    assert got["compressions"] == [f32(p["CriticalCompression"]) for p in g]
    assert got["stretches"] == [f32(p["CriticalStretch"]) for p in g]
    assert got["constitutive_models"] == [MODEL[p["ConstitutiveModel"]] for p in g]
    assert got["origin"] == [f32(v) for v in y["Grid"]["Origin"]]
    assert got["resolution"] == list(y["Grid"]["Resolution"])
    assert got["stride"] == f32(y["Grid"]["Stride"])
    assert got["friction_coeff"] == f32(y["Grid"]["BoundaryFrictionCoefficient"])
    assert got["blend_coeff"] == f32(y["Transfer"]["BlendCoefficient"])
    assert got["interp_type"] == INTERP[y["Interpolation"]["Type"]]
    assert got["delta_time"] == f32(y["Simulate"]["deltaTime"])
    assert got["offline"] == bool(y["Render"]["Offline"])
    assert got["particle_frag_shader_path"] == y["Render"]["ParticleFragmentShader"]
    assert abs(got["gravity"] - 9.80665) < 1e-6 and got["steps"] == 0


def test_parser_errors(tmp_path):
    base = open(os.path.join(CFG_DIR, "snow_cube.yaml")).read()
    bad_model = tmp_path / "bad_model.yaml"
    bad_model.write_text(base.replace("FixedCorotated", "Plasticine"))
    rc, text = _parse_cpp(str(bad_model))
    assert rc == 1 and "Unsupported constitutive model: Plasticine" in text
    bad_interp = tmp_path / "bad_interp.yaml"
    bad_interp.write_text(base.replace("CubicBSpline", "Quartic"))
    rc, text = _parse_cpp(str(bad_interp))
    assert rc == 1 and "Unsupported interpolation type" in text
    missing = tmp_path / "missing.yaml"
    missing.write_text(base.replace("  BoundaryFrictionCoefficient: 0.9\n", ""))
    rc, text = _parse_cpp(str(missing))
    assert rc == 1 and "BoundaryFrictionCoefficient" in text
    rc, text = _parse_cpp(str(tmp_path / "nope.yaml"))
    assert rc == 1 and "bad file" in text
    extra = tmp_path / "extra.yaml"
    extra.write_text(base.replace("  deltaTime: 1e-3", "  deltaTime: 1e-3\n  Gravity: 9.8  # optional\n  Steps: 250"))
    rc, text = _parse_cpp(str(extra))
    got = json.loads(text)
    assert rc == 0 and abs(got["gravity"] - 9.8) < 1e-6 and got["steps"] == 250


def test_capi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "mpm_hip.h")).read()
    declared = set(re.findall(r"MPM_API\s+[\w\s\*]+?\b(mpm_\w+)\s*\(", header))
    assert declared == set(N.SIGNATURES), declared ^ set(N.SIGNATURES)
    lib = N.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.mpm_version().startswith(b"mpm_hip")


def test_capi_rejects_bad_descriptors_without_gpu():
    """Validation happens before any device call."""
    from mpm_amd.solver import ParticleInput, SolverDesc
    lib = N.load()
    p = ParticleInput(np.array([[0.5, 0.5, 0.5]]), np.zeros((1, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    h = C.c_void_p()
    cases = [
        (SolverDesc(interp=N.LINEAR, transfer=N.APIC), N.MPM_ERR_UNSUPPORTED, "singular"),
        (SolverDesc(interp=7), N.MPM_ERR_INVALID, "interpolation"),
        (SolverDesc(resolution=(0, 4, 4)), N.MPM_ERR_INVALID, "resolution"),
        (SolverDesc(stride=0.0), N.MPM_ERR_INVALID, "stride"),
        (SolverDesc(dt=float("nan")), N.MPM_ERR_INVALID, "non-finite"),
    ]
    for d, code, msg in cases:
        cd = d.to_c()
        ci = p.to_c()
        rc = lib.mpm_create(C.byref(cd), C.byref(ci), C.byref(h))
        assert rc == code, (d, rc)
        assert msg in lib.mpm_last_error().decode()
        assert not h.value
    bad = ParticleInput(np.array([[0.5, 0.5, 0.5]]), np.zeros((1, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 5)
    cd = SolverDesc().to_c()
    ci = bad.to_c()
    assert lib.mpm_create(C.byref(cd), C.byref(ci), C.byref(h)) == N.MPM_ERR_INVALID
    assert "constitutive model" in lib.mpm_last_error().decode()


def test_scene_generators_are_seeded_and_sized():
    from mpm_amd import scenes
    a = scenes.snow_cube(seed=0)
    b = scenes.snow_cube(seed=0)
    assert a.particles.count == 16384
    assert np.array_equal(a.particles.position, b.particles.position)
    assert np.isclose(a.particles.mass[0], 400 * 0.2 ** 3 / 16384, rtol=1e-6)
    assert (a.particles.position >= 0.4).all() and (a.particles.position <= 0.6).all()
    tb = scenes.two_balls(seed=1)
    n = tb.particles.count
    assert 0.9 * 2 * 70300 < n < 1.1 * 2 * 70300  # ~70.3k per ball at 8 per cell, 128^3
    s = scenes.apic_slab(seed=2, nx=16, ny=8, nz=16, res=32)
    assert s.particles.count == 16 * 8 * 16
    assert (s.particles.young >= 1.4e5).all() and (s.particles.young <= 5.6e5).all()
    u = scenes.Rng(5).uniform(100000)
    assert 0.49 < u.mean() < 0.51 and u.min() >= 0 and u.max() < 1
    batch = scenes.uneven_batch(seed=3, count=8)
    assert all(2000 <= sc.particles.count <= 200000 for sc in batch)
