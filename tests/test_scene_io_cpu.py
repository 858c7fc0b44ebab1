"""Scene files: binary particle groups (.xyzb) read identically by the
Python loader and the C++ builder's reader, YAML round trips through both
parsers, the reference's own .xyz inputs, and the generator command line."""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.scene_io import load_yaml, read_group, read_xyzb, write_scene, write_xyzb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mpm_amd", "lib", "libchains_mpm.so")
REF_DATA = os.path.join(ROOT, "tests", "golden", "reference_inputs", "data")


def _facade():
    if not os.path.exists(LIB):
        pytest.skip("facade library not built")
    lib = C.CDLL(LIB)
    f = C.POINTER(C.c_float)
    lib.chains_read_group.argtypes = [C.c_char_p, C.c_float, f, f, f, f, C.c_int64, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int)]
    lib.chains_parse_yaml_json.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    return lib


def _cread(path, stride, expect_rc=0):
    lib = _facade()
    cnt, fields = C.c_int64(0), C.c_int(0)
    rc = lib.chains_read_group(path.encode(), stride, None, None, None, None, 0, C.byref(cnt), C.byref(fields))
    assert rc == expect_rc
    n = cnt.value
    out = {"position": np.zeros((n, 3), np.float32), "velocity": np.zeros((n, 3), np.float32),
           "young": np.zeros(n, np.float32), "mass": np.zeros(n, np.float32)}
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.chains_read_group(path.encode(), stride, fp(out["position"]), fp(out["velocity"]), fp(out["young"]),
                                 fp(out["mass"]), n, C.byref(cnt), C.byref(fields)) == expect_rc
    return out, fields.value


# Text groups the reference's reader (src/mpm_solver3d_builder.cu:10-30, its
# result ignored at :40) turns into these points: every line is read with
# three float extractions, the first line that fails stops the file, and the
# points before it are kept.
TEXT_CASES = [
    ("trailing_blank_line", "1 2 3\n4 5 6\n\n", [[1, 2, 3], [4, 5, 6]], False),
    ("bad_line_mid_file", "1 2 3\n4 5\n7 8 9\n", [[1, 2, 3]], False),
    ("crlf_and_extra_text", "1 2 3\r\n-.5 +2e1 3.25 tail\r\n", [[1, 2, 3], [-0.5, 20, 3.25]], True),
    ("glued_signs", "1-2 3\n", [[1, -2, 3]], True),
    ("no_final_newline", "\t0.125 0.25 0.5", [[0.125, 0.25, 0.5]], True),
    ("leading_blank_line", "\n1 2 3\n", [], False),
    ("letters", "1 2 x\n", [], False),
    ("empty_file", "", [], True),
]


@pytest.mark.parametrize("name,text,points,complete", TEXT_CASES, ids=[c[0] for c in TEXT_CASES])
def test_text_group_prefix_rule(tmp_path, name, text, points, complete):
    """The C++ builder's reader, the Python loader and the reference's
    observable result agree, including the prefix kept before a bad line."""
    from mpm_amd.scene_io import read_text_group
    p = str(tmp_path / (name + ".xyz"))
    with open(p, "wb") as f:
        f.write(text.encode())
    want = np.array(points, np.float32).reshape(-1, 3)
    pts, ok = read_text_group(p)
    assert ok == complete and np.array_equal(pts, want), (pts, ok)
    if len(points) == 0 and not complete:
        _cread(p, 1.0, expect_rc=-1)  # nothing before the bad line: an empty group
        return
    cc, _ = _cread(p, 1.0, expect_rc=0 if complete else 1)
    assert np.array_equal(cc["position"], want)
    assert np.array_equal(read_group(p, 1.0)["position"], want)


@pytest.mark.parametrize("world", [True, False])
@pytest.mark.parametrize("extra", [False, True])
def test_xyzb_python_and_cpp_readers_agree(tmp_path, world, extra):
    rng = np.random.default_rng(1)
    n = 1000
    pos = rng.uniform(0, 50, (n, 3)).astype(np.float32)
    vel = rng.normal(0, 1, (n, 3)).astype(np.float32) if extra else None
    young = rng.uniform(1e5, 4e5, n).astype(np.float32) if extra else None
    mass = rng.uniform(1e-4, 2e-4, n).astype(np.float32) if extra else None
    p = str(tmp_path / "g.xyzb")
    write_xyzb(p, pos, vel, young, mass, world=world)
    raw = read_xyzb(p)
    assert np.array_equal(raw["position"], pos) and raw["world"] == world
    stride = 1.0 / 96
    py = read_group(p, stride)
    cc, fields = _cread(p, stride)
    assert np.array_equal(py["position"], cc["position"])
    want = pos if world else np.float32(stride) * pos
    assert np.array_equal(py["position"], want)
    assert fields == (7 if extra else 0)
    if extra:
        for k, v in (("velocity", vel), ("young", young), ("mass", mass)):
            assert np.array_equal(py[k], v) and np.array_equal(cc[k], v)


def test_malformed_xyzb_rejected(tmp_path):
    p = str(tmp_path / "bad.xyzb")
    write_xyzb(p, np.zeros((10, 3), np.float32))
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-4])  # truncated
    with pytest.raises(ValueError):
        read_xyzb(p)
    lib = _facade()
    cnt, fields = C.c_int64(0), C.c_int(0)
    assert lib.chains_read_group(p.encode(), 1.0, None, None, None, None, 0, C.byref(cnt), C.byref(fields)) == -1
    open(p, "wb").write(raw + b"\0\0\0\0")  # trailing bytes
    assert lib.chains_read_group(p.encode(), 1.0, None, None, None, None, 0, C.byref(cnt), C.byref(fields)) == -1


def test_reference_xyz_text_readers_agree():
    # the reference's own input (data/cube_sparse.xyz, 11,907 points)
    p = os.path.join(REF_DATA, "cube_sparse.xyz")
    stride = float(np.float32(5e-3))
    py = read_group(p, stride)
    cc, fields = _cread(p, stride)
    assert py["position"].shape == (11907, 3) and fields == 0
    np.testing.assert_array_max_ulp(py["position"], cc["position"], maxulp=1)


@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_yaml_round_trip_both_parsers(tmp_path, transfer):
    sc = scenes.snow_cube(seed=3, n=2000, res=32, transfer=transfer, velocity=(0.1, 0.2, -0.3))
    sc.particles.young[:] = np.linspace(1e5, 4e5, 2000, dtype=np.float32)
    y = write_scene(sc.desc, sc.particles, str(tmp_path), "cube", steps=7)
    desc, parts = load_yaml(y, base_dir=str(tmp_path))
    # the descriptor as the device sees it (float32 fields)
    for f in ("origin", "stride", "friction", "blend", "dt", "gravity"):
        assert np.array_equal(np.float32(getattr(desc, f)), np.float32(getattr(sc.desc, f))), f
    assert (desc.resolution, desc.interp, desc.transfer) == (sc.desc.resolution, sc.desc.interp, sc.desc.transfer)
    for f in ("position", "velocity", "mass", "hardening", "young", "poisson", "critical_compression",
              "critical_stretch", "model"):
        assert np.array_equal(getattr(parts, f), getattr(sc.particles, f)), f
    lib = _facade()
    buf = C.create_string_buffer(1 << 16)
    assert lib.chains_parse_yaml_json(y.encode(), buf, len(buf)) == 0, buf.value
    cfg = json.loads(buf.value.decode())
    assert cfg["scheme"] == (1 if transfer == N.APIC else 0) and cfg["steps"] == 7
    assert cfg["group_paths"] == ["cube.xyzb"]
    assert np.float32(cfg["stride"]) == np.float32(sc.desc.stride)
    assert np.float32(cfg["delta_time"]) == np.float32(sc.desc.dt)
    assert np.float32(cfg["gravity"]) == np.float32(sc.desc.gravity)


def test_generator_cli(tmp_path):
    out = str(tmp_path / "cube")
    r = subprocess.run([sys.executable, "-m", "mpm_amd.scene_io", "--workload", "snow_cube", "--out", out],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    desc, parts = load_yaml(os.path.join(out, "snow_cube.yaml"), base_dir=out)
    sc = scenes.snow_cube()
    assert parts.count == sc.particles.count and desc.resolution == sc.desc.resolution
    assert np.array_equal(parts.position, sc.particles.position)
    r = subprocess.run([sys.executable, "-m", "mpm_amd.scene_io", "--workload", "uneven_batch", "--count", "3",
                        "--out", str(tmp_path / "batch")], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    index = json.load(open(tmp_path / "batch" / "batch.json"))
    assert len(index) == 3 and all(os.path.exists(tmp_path / "batch" / e["yaml"]) for e in index)
