"""Build the native libraries in-tree.

* ``mpm_amd/lib/libmpm_hip.so`` -- HIP kernels + C-ABI (hipcc, gfx950)
* ``mpm_amd/lib/libchains_mpm.so`` -- C++ drop-in facade (namespace chains)
* ``mpm_amd/lib/mpm_headless`` -- headless driver of the facade

Usage: ``python -m mpm_amd.build [--jobs N] [--verbose]``.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FACADE = os.path.join(PKG, "facade")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["sort.hip", "kernels.hip", "capi.hip", "batch.hip", "incsort.hip"]
FACADE_SOURCES = ["mpm3d_parser.cpp", "material_point3d.cpp", "mpm_solver3d.cpp", "mpm_solver3d_builder.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPM_OFFLOAD_ARCH", "gfx950")
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 generation.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I", INCLUDE,
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INCLUDE, "-I", FACADE]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    if verbose and r.stdout.strip():
        print(r.stdout)
    return r.stdout


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build_hip(jobs=4, verbose=False, force=True):
    os.makedirs(OBJ, exist_ok=True)
    headers = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    headers.append(os.path.join(INCLUDE, "mpm_hip.h"))
    objs = []
    tasks = []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _stale(o, [s] + headers):
            tasks.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), tasks))
    out = os.path.join(LIB, "libmpm_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, out)
    return out


def build_facade(verbose=False):
    srcs = [os.path.join(FACADE, s) for s in FACADE_SOURCES if os.path.exists(os.path.join(FACADE, s))]
    if not srcs:
        return None
    out = os.path.join(LIB, "libchains_mpm.so")
    tmp = out + ".tmp"
    _run([CXX] + CXX_FLAGS + ["-shared", "-o", tmp] + srcs + ["-L", LIB, "-lmpm_hip", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, out)
    drv = os.path.join(FACADE, "mpm_headless.cpp")
    if os.path.exists(drv):
        exe = os.path.join(LIB, "mpm_headless")
        _run([CXX] + CXX_FLAGS + ["-o", exe + ".tmp", drv, "-L", LIB, "-lchains_mpm", "-lmpm_hip",
                                  "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(exe + ".tmp", exe)
    return out


def build_all(jobs=4, verbose=False):
    lib = build_hip(jobs=jobs, verbose=verbose)
    fac = build_facade(verbose=verbose)
    if os.path.isdir(OBJ):
        shutil.rmtree(OBJ, ignore_errors=True)
    return lib, fac


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    lib, fac = build_all(jobs=a.jobs, verbose=a.verbose)
    print("built", lib, fac or "")
    return 0


if __name__ == "__main__":
    sys.exit(main())
