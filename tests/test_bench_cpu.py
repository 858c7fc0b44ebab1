"""bench.py pieces that need no GPU: the single-GPU refusal, the byte model
the roofline uses, and the CPU-baseline legs on a small scene."""
from __future__ import annotations

import os
import subprocess
import sys

import bench
from mpm_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_byte_model():
    # APIC 364 B and PIC-FLIP 304 B per particle-update, 56 B per active node
    assert bench.BYTES[1]["p2g"] + bench.BYTES[1]["g2p"] == 364
    assert bench.BYTES[0]["p2g"] + bench.BYTES[0]["g2p"] == 304
    assert bench.GRID_BYTES["p2g"] + bench.GRID_BYTES["g2p"] == 56
    # the fused kernel is credited with both transfers
    assert bench.BYTES[1]["g2p2g"] == 364 and bench.GRID_BYTES["g2p2g"] == 56


def test_cpu_baseline_legs_small_scene():
    sc = scenes.snow_cube(seed=0, n=2000, res=32)
    serial = bench.cpu_baseline_serial(sc, seconds_target=0.5)
    assert serial["cores"] == 1 and serial["kind"] == "port" and serial["value"] > 0
    assert "whole workload" in serial["sample"]
    allc = bench.cpu_baseline(sc, seconds_target=0.5)
    assert allc["cores"] >= 1 and allc["value"] > 0 and allc["serial"]["value"] > 0
