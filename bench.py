"""Benchmark: explicit MPM step throughput on one MI355X.

Workload (default): the large APIC slab -- 256x128x256 = 8,388,608 particles
on a jittered lattice in [0.1,0.9]x[0.1,0.5]x[0.1,0.9], 256^3 grid,
quadratic B-spline, APIC, v ~ N(0, 0.5^2), E = 1.4e5 * U[1,4], dt = 2.5e-5,
FP32, seeded synthetic data.  A step is one simulateOneStep (sort, P2G,
grid update, G2P + F update + advection) over all particles; particles and
grid are resident in HBM before the timed region starts.

Prints ONE JSON line.  ``value`` is particle-updates/s over the timed steps.
``roofline`` is for the dominant kernel (largest total time), with
algorithmic bytes per launch (reads/writes of particle state and active grid
nodes, per the byte model below) divided by its mean HIP-event duration.
``cpu_baseline`` is the C++ oracle (a port of the reference step) on the
host: on all host cores (OpenMP loops, atomic grid sums, as the reference's
atomicAdd scatter) over the whole workload for two steps, with the serial
oracle on a sub-box of it beside that (``serial``).  ``small_scene`` times the reference's
CPU-runnable configuration (16,384-particle snow cube, PIC-FLIP) eagerly and
as a replayed hipGraph, beside the serial oracle on the same scene.

Algorithmic bytes (FP32): APIC per particle-update 364 B = P2G 148 B
(x, v, B, Fe, Fp, m, V0, mu0, lambda0) + G2P 216 B (reads x, Fe, Fp; writes
x, v, B, Fe, Fp); PIC-FLIP 304 B = 112 + 192.  Grid: 28 B per active node
written by P2G and 28 B read by G2P (active = mass > 0, counted per step).
The fused APIC kernel (g2p2g: one step's G2P and the next step's P2G) is
credited with both: 364 B per particle and 56 B per active node.

``roofline.traffic`` is the kernel's DRAM bytes per launch from the L2's
request-size-resolved counters (TCC_EA0_RDREQ_DRAM_32B and
TCC_EA0_WRREQ_WRITE_DRAM_32B / _ATOMIC_DRAM_32B, in 32-byte units: a 64-
or 128-byte request counts 2 or 4), checked on a plain 1 GiB device copy;
``traffic_detail`` also carries FETCH_SIZE / WRITE_SIZE scaled by what they
read for that copy.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12
METRIC = "particle-updates/sec & %HBM roofline; batched p50/p99 step latency vs plain"

BYTES = {  # per particle-update, per kernel
    1: {"p2g": 148, "g2p": 216, "g2p2g": 364},   # APIC
    0: {"p2g": 112, "g2p": 192},                 # PIC-FLIP
}
GRID_BYTES = {"p2g": 28, "g2p": 28, "g2p2g": 56}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, seconds_target=15.0):
    """All-cores oracle over the whole workload (2 timed steps after 1
    warm-up), and the serial oracle on a spatial sub-box of it (full grid)."""
    from oracle import pyoracle

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    o = pyoracle.OracleSolver(scene.desc, scene.particles, 32, threads=threads)
    o.time_steps(1)  # warm-up step
    t_all = o.time_steps(2)
    o.close()
    serial = cpu_baseline_serial(scene, seconds_target)
    return {
        "value": scene.particles.count * 2 / t_all,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "cpu": _cpu_model(),
        "sample": f"C++ oracle (FP32 restatement of the reference step) on {threads} threads, OpenMP over "
                  f"particles and nodes with atomic grid sums, the whole {scene.particles.count}-particle "
                  f"workload, 2 steps after 1 warm-up, {t_all:.1f} s",
        "serial": serial,
    }


def cpu_baseline_serial(scene, seconds_target=15.0):
    """Serial oracle on a spatial sub-box of the same workload (full grid)."""
    from oracle import pyoracle

    p = scene.particles
    x = p.position
    # x,z in the low corner of the domain: ~1/16 of the slab's particles
    # (scenes with few particles there run whole)
    sel = np.nonzero((x[:, 0] < 0.3) & (x[:, 2] < 0.3))[0]
    whole = sel.size < min(p.count, 100_000)
    sub = p if whole else p.subset(sel)
    where = "the whole workload" if whole else "sub-box x,z<0.3"
    o = pyoracle.OracleSolver(scene.desc, sub, 32)
    o.time_steps(1)  # warm-up step
    steps = 0
    t = 0.0
    while t < seconds_target and steps < 3:
        t += o.time_steps(1)
        steps += 1
    o.close()
    return {
        "value": sub.count * steps / t,
        "unit": "particle-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"serial C++ oracle (FP32 restatement of the reference step), {sub.count} particles "
                  f"({where}) on the full {'x'.join(str(r) for r in scene.desc.resolution)} grid, {steps} steps after "
                  f"1 warm-up, {t:.1f} s",
    }


def density_bench(solver, sc, frames=5, cpu=True):
    """Density raster (the writeToOpenVDB replacement) after each of `frames`
    steps: HIP-event time of the raster (sort of the current positions +
    mass-only P2G into its own plane), wall time of the block readback, and
    the reference's host raster (the oracle's restatement) on a sample.
    Algorithmic bytes: x and m per particle (16 B) + 4 B per nonzero node."""
    from mpm_amd import native as N

    n = sc.particles.count
    # the standalone raster: after a fused step the grid already holds the
    # mass sums of the current positions and the raster reads them, so the
    # raster's own work is timed with the two-kernel step
    solver.set_transfer_fusion(False)
    solver.profile(True)
    solver.profile_reset()
    for _ in range(frames):
        solver.step(1)
        solver.density_compute()
    solver.synchronize()
    prof = solver.profile_read()
    solver.profile(False)
    ms = prof["density"][1] / max(prof["density"][0], 1)
    solver.step(1)
    solver.synchronize()
    q0 = time.perf_counter()
    blk = solver.density_blocks()
    read_ms = 1e3 * (time.perf_counter() - q0)
    solver.set_transfer_fusion(True)
    solver.step(2)  # fused steps: the raster is the grid's mass plane
    solver.synchronize()
    q0 = time.perf_counter()
    blk_f = solver.density_blocks()
    read_fused_ms = 1e3 * (time.perf_counter() - q0)
    nz_f = int(np.count_nonzero(blk_f.values))
    nz = int(np.count_nonzero(blk.values))
    alg = 16 * n + 4 * nz
    out = {"frames": frames, "ms": ms, "particles_per_s": n / (ms * 1e-3), "blocks": blk.count,
           "nonzero_nodes": nz, "algorithmic_bytes": alg, "achieved_GBps": alg / (ms * 1e-3) / 1e9,
           "frac": alg / (ms * 1e-3) / HBM_PEAK, "block_readback_ms": read_ms,
           "after_fused_step": {"raster_ms": 0.0, "block_readback_ms": read_fused_ms, "blocks": blk_f.count,
                                "nonzero_nodes": nz_f},
           "parts_ms": {k: v[1] / max(v[0], 1) for k, v in prof.items()}}
    if cpu:
        from oracle import pyoracle

        x = sc.particles.position
        sel = np.nonzero((x[:, 0] < 0.3) & (x[:, 2] < 0.3))[0]
        t = pyoracle.time_density(sc.desc, x[sel], sc.particles.mass[sel])
        out["cpu_host_raster"] = {"particles_per_s": sel.size / t, "cores": 1, "kind": "port",
                                  "sample": f"serial host raster of {sel.size} particles (slab sub-box x,z<0.3), "
                                            f"{t:.2f} s"}
    fused = sc.desc.transfer == N.APIC and sc.desc.interp == N.QUADRATIC_BSPLINE
    log(f"density raster: {ms:.3f} ms/frame ({out['particles_per_s']:.3e} particles/s), {blk.count} blocks, "
        f"block readback {read_ms:.1f} ms" +
        (f"; after a fused step no raster pass, block readback {read_fused_ms:.1f} ms" if fused else ""))
    if not fused:
        out.pop("after_fused_step")
    return out


def _pmc_pass(counters, args, outdir, tag):
    """One rocprofv3 counter pass (counters: a space-separated list that fits
    one pass) over a short run of this script (child process group, killed
    after 240 s); returns {kernel name: {counter: mean value per dispatch}}."""
    import csv
    import shutil
    import signal
    import subprocess

    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    names = counters.split()
    d = os.path.join(outdir, f"{tag}_{names[0]}")
    cmd = [prof, "--pmc"] + names + ["--output-format", "csv", "-d", d, "-o", tag, "--",
                                     sys.executable, os.path.abspath(__file__)] + args
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, env=env, start_new_session=True)
    try:
        p.wait(timeout=240)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        return None
    sums, disp = {}, {}
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for r in csv.DictReader(open(os.path.join(root, f))):
                    k = r["Kernel_Name"]
                    c = sums.setdefault(k, {})
                    c[r["Counter_Name"]] = c.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
                    disp.setdefault(k, set()).add(r.get("Dispatch_Id", r.get("Correlation_Id", "")))
    return {k: {c: v / max(len(disp[k]), 1) for c, v in cs.items()} for k, cs in sums.items()} or None


DRAM_COUNTERS = ["TCC_EA0_RDREQ_DRAM_32B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum",
                 "TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum"]


def pmc_traffic(workload, seed_args, outdir):
    """DRAM bytes per launch of the step's item kernels (P2G and G2P, or the
    fused g2p2g) over a 3-step run: the L2's request-size-resolved DRAM
    counters (32-byte units) in one pass, checked on a plain 1 GiB device
    copy; beside them FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
    scaled by what it reads for that copy."""
    run = ["--workload", workload, "--steps", "3", "--warmup", "1", "--no-cpu-baseline", "--batch-scenes", "0",
           "--serve-seconds", "0", "--no-small", "--no-traffic"] + seed_args
    copy = ["--calibrate-copy", "1024"]
    dram, cdram = _pmc_pass(" ".join(DRAM_COUNTERS), run, outdir, "step"), \
        _pmc_pass(" ".join(DRAM_COUNTERS), copy, outdir, "copy")
    fe, wr = _pmc_pass("FETCH_SIZE", run, outdir, "step"), _pmc_pass("WRITE_SIZE", run, outdir, "step")
    cfe, cwr = _pmc_pass("FETCH_SIZE", copy, outdir, "copy"), _pmc_pass("WRITE_SIZE", copy, outdir, "copy")
    if not (dram and cdram):
        return None
    gib = float(1 << 30)

    def rw(table, k):  # (read, write) bytes from the DRAM counters of kernel k
        v = table[k]
        return (32.0 * v.get(DRAM_COUNTERS[0], 0.0),
                32.0 * (v.get(DRAM_COUNTERS[1], 0.0) + v.get(DRAM_COUNTERS[2], 0.0)))

    ck = [k for k in cdram if "copyBuffer" in k]
    out = {"method": "TCC_EA0_*_DRAM_32B x 32 B"}
    if ck:
        r, w = rw(cdram, ck[0])
        out["copy_check"] = {"read_bytes": r, "write_bytes": w, "read_frac": r / gib, "write_frac": w / gib}
    sf = sw = None
    if fe and wr and cfe and cwr:
        cks = [k for k in cfe if "copyBuffer" in k]
        if cks and cks[0] in cwr:
            sf = 1024 * 1024 / cfe[cks[0]]["FETCH_SIZE"]
            sw = 1024 * 1024 / cwr[cks[0]]["WRITE_SIZE"]
            out.update({"fetch_scale": sf, "write_scale": sw})
    # the step's instances: P2G mode 0 (7 planes) or 2 (folded APIC grid),
    # G2P source 0 or 3, the fused kernel; not initialize()'s mass / volume
    # passes
    step_kernels = {"p2g": ("k_p2g<", (", 0, false>", ", 2, false>")),
                    "g2p": ("k_g2p<", (", 0, false,", ", 3, false,")),
                    "g2p2g": ("k_g2p2g<", ("",))}
    for name, (key, tags) in step_kernels.items():
        ks = [k for k in dram if key in k and any(t in k for t in tags)]
        if not ks:
            continue
        k = max(ks, key=lambda q: rw(dram, q)[0])
        r, w = rw(dram, k)
        out[name] = {"kernel": k.split("(")[0], "read_bytes": r, "write_bytes": w, "bytes": r + w}
        if sf and k in fe and k in wr:
            f_kib, w_kib = fe[k]["FETCH_SIZE"], wr[k]["WRITE_SIZE"]
            out[name].update({"fetch_kib": f_kib, "write_kib": w_kib,
                              "scaled_bytes": (f_kib * sf + w_kib * sw) * 1024.0})
    return out


def batch_bench(count, steps, warmup, seed=3):
    """Uneven scene batch: the same scenes stepped one at a time (the plain
    simulateOneStep loop) and as one batch (eager launches and hipGraph
    replay, launch parameters from the autotuner)."""
    from mpm_amd import native as N
    from mpm_amd import scenes
    from mpm_amd.batch import Batch
    from mpm_amd.solver import Solver

    t0 = time.time()
    scs = scenes.uneven_batch(seed=seed, count=count)
    ntot = sum(s.particles.count for s in scs)
    log(f"batch workload: {count} scenes, {ntot} particles, generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    solvers = [Solver(s.desc, s.particles) for s in scs]
    for s in solvers:
        s.synchronize()
    log(f"batch scenes created in {time.time() - t0:.1f}s")
    # plain loop: every scene's own step, one scene after another
    for _ in range(warmup):
        for s in solvers:
            s.step(1)
    for s in solvers:
        s.synchronize()
    rounds = []
    for _ in range(steps):
        q0 = time.perf_counter()
        for s in solvers:
            s.step(1)
        for s in solvers:
            s.synchronize()
        rounds.append(1e3 * (time.perf_counter() - q0))
    rounds = np.array(rounds)
    plain = ntot * steps / (rounds.sum() * 1e-3)
    # one scene at a time, synchronised: that scene's own step latency
    solo = []
    for s in solvers[: min(64, count)]:
        q0 = time.perf_counter()
        s.step(1)
        s.synchronize()
        solo.append(1e3 * (time.perf_counter() - q0))
    log(f"plain loop: {rounds.mean():.2f} ms per round of {count} scenes, {plain:.3e} p-u/s")
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    for s in solvers:
        b.add(s)
    tune = b.autotune(trial_steps=4)
    log(f"autotune: item_max={tune.item_max} g2p_threads={tune.g2p_threads} workgroups={tune.workgroups} "
        f"fuse_transfer={tune.fuse_transfer} ({tune.step_ms:.3f} ms/step)")
    out = {}
    for mode in ("eager", "graph"):
        run = b.step if mode == "eager" else b.step_graph
        run(warmup)
        b.synchronize()
        # throughput: back-to-back steps with no events in the stream
        q0 = time.perf_counter()
        run(steps)
        b.synchronize()
        el = time.perf_counter() - q0
        # per-step latency: a second pass with a hipEvent pair around each step
        b.timing(True)
        run(steps)
        b.synchronize()
        t = b.read_times()
        b.timing(False)
        out[mode] = {"value": ntot * steps / el, "ms_per_step": 1e3 * el / steps,
                     "step_ms_p50": float(np.percentile(t, 50)), "step_ms_p99": float(np.percentile(t, 99))}
        log(f"batched ({mode}): {1e3 * el / steps:.2f} ms/step, {ntot * steps / el:.3e} p-u/s, "
            f"step p50/p99 {out[mode]['step_ms_p50']:.2f}/{out[mode]['step_ms_p99']:.2f} ms")
    info = b.info()
    b.close()
    for s in solvers:
        s.close()
    best = max(out.values(), key=lambda d: d["value"])
    return {
        "workload": f"uneven_batch: {count} scenes, N log-uniform [2k,200k], grids 32^3/64^3/128^3, APIC quadratic",
        "scenes": count, "particles": ntot, "steps": steps,
        "plain": {"value": plain, "round_ms_p50": float(np.percentile(rounds, 50)),
                  "round_ms_p99": float(np.percentile(rounds, 99)),
                  "scene_step_ms_p50": float(np.percentile(solo, 50)),
                  "scene_step_ms_p99": float(np.percentile(solo, 99))},
        "batched": out["eager"], "batched_graph": out["graph"],
        "speedup_vs_plain": best["value"] / plain,
        "tune": {"item_max": tune.item_max, "g2p_threads": tune.g2p_threads, "workgroups": tune.workgroups,
                 "fuse_transfer": tune.fuse_transfer, "bucket": tune.bucket},
        "work_items": info["work_items"], "key_bits": info["key_bits"],
    }


def small_scene_bench(steps=200, cpu=True):
    """The reference's CPU-runnable configuration (snow cube, 16,384
    particles, 64^3, PIC-FLIP quadratic): launch-bound on the GPU, so it is
    timed eagerly and as a replayed hipGraph, beside the serial oracle."""
    from mpm_amd import scenes
    from mpm_amd.solver import Solver

    sc = scenes.workload("snow_cube")
    n = sc.particles.count
    out = {"workload": f"snow_cube: {n} particles, 64^3 grid, PIC-FLIP quadratic, dt 1e-4", "steps": steps}
    for mode in ("eager", "graph"):
        s = Solver(sc.desc, sc.particles)
        run = s.step if mode == "eager" else s.step_graph
        run(5)
        s.synchronize()
        t0 = time.perf_counter()
        run(steps)
        s.synchronize()
        el = time.perf_counter() - t0
        s.close()
        out[mode] = {"ms_per_step": 1e3 * el / steps, "value": n * steps / el}
    if cpu:
        from oracle import pyoracle

        o = pyoracle.OracleSolver(sc.desc, sc.particles, 32)
        o.time_steps(1)
        t = o.time_steps(10)
        o.close()
        out["cpu_serial"] = {"ms_per_step": 1e3 * t / 10, "value": n * 10 / t, "cores": 1, "kind": "port"}
    log(f"small scene: eager {out['eager']['ms_per_step']:.3f} ms/step, graph {out['graph']['ms_per_step']:.3f} ms/step"
        + (f", serial oracle {out['cpu_serial']['ms_per_step']:.1f} ms/step" if cpu else ""))
    return out


def serve_plain(reqs):
    """The same requests one at a time, in arrival order, as a plain caller
    runs them: create the scene, step it to the end, read the result back,
    free it (no batching, no overlap).  Returns (wall s, particle-updates)."""
    from mpm_amd.solver import Solver

    t0 = time.perf_counter()
    pu = 0
    for i, r in enumerate(reqs):
        s = Solver(r.desc, r.particles)
        s.step(r.steps)
        s.read_particles()
        s.close()
        pu += r.particles.count * r.steps
        if (i + 1) % 100 == 0:
            log(f"plain: {i + 1}/{len(reqs)} requests in {time.perf_counter() - t0:.1f}s")
    return time.perf_counter() - t0, pu


def serve_bench(window_s, seed=4, rate=20.0, plain=False, saturate=0):
    """Poisson request stream (seeded) over a window of window_s seconds,
    served to completion under a 200 GB HBM budget.  plain: the same stream
    also run one request at a time (served / plain).  saturate: the first
    `saturate` requests of the stream also served arriving all at once (the
    server's capacity)."""
    import dataclasses

    from mpm_amd import scenes
    from mpm_amd.serve import Server, requests_from_scenes

    t0 = time.time()
    reqs = requests_from_scenes(scenes.request_stream(seed=seed, rate=rate, duration=window_s))
    log(f"serve workload: {len(reqs)} requests over {window_s}s, generated in {time.time() - t0:.1f}s")
    srv = Server(budget_bytes=200e9, tick_steps=16)
    try:
        rep = srv.run(reqs)
    finally:
        srv.close()
    d = rep.summary()
    log(f"served {rep.completed} requests in {rep.wall_s:.1f}s: {rep.particle_updates_per_s:.3e} p-u/s, "
        f"request p50/p99 {rep.request_s_p50:.2f}/{rep.request_s_p99:.2f}s")
    d["workload"] = (f"request_stream: Poisson {rate}/s for {window_s}s, N log-uniform [5k,500k], "
                     f"steps U[200,2000], PIC-FLIP/APIC 50/50, 64^3-256^3, 200 GB budget")
    if saturate > 0:
        burst = [dataclasses.replace(r, arrival=0.0) for r in reqs[:saturate]]
        srv = Server(budget_bytes=200e9, tick_steps=16)
        try:
            sat = srv.run(burst)
        finally:
            srv.close()
        d["saturated"] = {"requests": sat.completed, "wall_s": sat.wall_s,
                          "particle_updates_per_s": sat.particle_updates_per_s,
                          "step_ms_p50": sat.step_ms_p50, "step_ms_p99": sat.step_ms_p99,
                          "peak_scenes": sat.peak_scenes, "peak_bytes": sat.peak_bytes}
        log(f"saturated: {sat.completed} requests arriving at once, {sat.wall_s:.1f}s, "
            f"{sat.particle_updates_per_s:.3e} p-u/s, peak {sat.peak_scenes} scenes")
    if plain:
        wall, pu = serve_plain(reqs)
        d["plain"] = {"wall_s": wall, "particle_updates": pu, "particle_updates_per_s": pu / wall}
        d["served_vs_plain"] = rep.particle_updates_per_s / (pu / wall)
        log(f"plain one-at-a-time: {wall:.1f}s, {pu / wall:.3e} p-u/s; served/plain {d['served_vs_plain']:.2f}x")
    return d


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="apic_slab", choices=["apic_slab", "two_balls", "snow_cube"])
    ap.add_argument("--graph", action="store_true", help="replay the step as a captured hipGraph")
    ap.add_argument("--no-fusion", action="store_true",
                    help="APIC quadratic: G2P and the next P2G as two kernels instead of the fused one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the 16k-particle snow cube timing")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--batch-scenes", type=int, default=256, help="scenes of the batched run (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="request arrivals per second")
    ap.add_argument("--serve-plain", action="store_true",
                    help="also run the served stream one request at a time (served / plain)")
    ap.add_argument("--serve-saturate", type=int, default=0, metavar="N",
                    help="also serve the stream's first N requests arriving at once (capacity)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the served request stream (0: skip)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes that fill roofline.traffic")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    a = ap.parse_args(argv)

    if a.calibrate_copy:
        import torch

        n = a.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes_each_way": n * 4, "launches": 3}), flush=True)
        return 0

    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the solver shards nothing across GPUs",
              file=sys.stderr)
        return 2

    from mpm_amd import scenes
    from mpm_amd.solver import Solver

    # counter passes first, in child processes, before this process uses the GPU
    traffic = None
    if not a.no_traffic:
        t0 = time.time()
        import shutil
        import tempfile

        outdir = tempfile.mkdtemp(prefix="mpm_bench_pmc_")
        try:
            extra = ([] if a.seed is None else ["--seed", str(a.seed)]) + (["--no-fusion"] if a.no_fusion else [])
            traffic = pmc_traffic(a.workload, extra, outdir)
        finally:
            shutil.rmtree(outdir, ignore_errors=True)
        log(f"pmc traffic passes in {time.time() - t0:.1f}s: "
            + (", ".join(f"{k}={v['bytes'] / 1e9:.3f} GB" for k, v in traffic.items()
                          if isinstance(v, dict) and "bytes" in v)
               if traffic else "unavailable"))

    kw = {} if a.seed is None else {"seed": a.seed}
    t0 = time.time()
    sc = scenes.workload(a.workload, **kw)
    n = sc.particles.count
    log(f"workload {a.workload}: {n} particles, grid {sc.desc.resolution}, generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    solver = Solver(sc.desc, sc.particles)
    if a.no_fusion:
        solver.set_transfer_fusion(False)
    solver.synchronize()
    log(f"created + initialized in {time.time() - t0:.1f}s")

    run = solver.step_graph if a.graph else solver.step
    for _ in range(a.warmup):
        run(1)
    solver.synchronize()
    log(f"warmup {a.warmup} steps done")

    # timed region: K steps; HIP events on the solver stream around the item
    # kernels (p2g, g2p) only: events around every launch group would add
    # ~0.1 ms of stream gaps per step
    if not a.graph:
        solver.profile(2)
        solver.profile_reset()
    s0 = solver.stats()
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run(1)
    solver.synchronize()
    t1 = time.perf_counter()
    s1 = solver.stats()
    prof = solver.profile_read() if not a.graph else {}
    solver.profile(False)
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / a.steps
    value = n * a.steps / elapsed
    act = (s1["active_nodes_sum"] - s0["active_nodes_sum"]) / a.steps
    log(f"timed {a.steps} steps: {ms_per_step:.3f} ms/step, {value:.3e} p-u/s, mean active nodes {act:.0f}")

    apic = sc.desc.transfer == 1
    bp = BYTES[1 if apic else 0]
    step_bytes = n * (bp["p2g"] + bp["g2p"]) + act * (GRID_BYTES["p2g"] + GRID_BYTES["g2p"])
    step_frac = step_bytes / (elapsed / a.steps) / HBM_PEAK

    # per-kernel breakdown of the whole step: a separate pass with events
    # around every launch group (p2g and g2p keep their timed-region figures)
    kernels = {}
    if prof:
        solver.profile(1)
        solver.profile_reset()
        nb = min(50, a.steps)
        for _ in range(nb):
            run(1)
        solver.synchronize()
        for k, (launches, tot) in solver.profile_read().items():
            kernels[k] = {"launches": launches, "mean_ms": tot / max(launches, 1), "pass": f"breakdown ({nb} steps)"}
        solver.profile(False)
        for k, (launches, tot) in prof.items():
            kernels[k] = {"launches": launches, "mean_ms": tot / max(launches, 1), "pass": "timed"}

    roof = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1][1])[0]
        mean_s = kernels[dom]["mean_ms"] * 1e-3
        if dom in bp:
            alg = n * bp[dom] + act * GRID_BYTES[dom]
            tr = traffic.get(dom) if traffic else None
            roof = {"bound": "hbm", "kernel": dom, "achieved": alg / mean_s / 1e9, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": alg / mean_s / HBM_PEAK, "traffic": tr["bytes"] if tr else None,
                    "algorithmic_bytes": alg, "mean_ms": kernels[dom]["mean_ms"],
                    "traffic_detail": traffic}
        else:
            roof = {"bound": "hbm", "kernel": dom, "achieved": None, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": None, "traffic": None, "mean_ms": kernels[dom]["mean_ms"]}
        log("kernels: " + ", ".join(f"{k}={v['mean_ms']:.3f}ms x{v['launches']}" for k, v in kernels.items()))

    # per-step latency (synchronised steps)
    lat = []
    for _ in range(min(50, a.steps)):
        q0 = time.perf_counter()
        run(1)
        solver.synchronize()
        lat.append(1e3 * (time.perf_counter() - q0))
    lat = np.array(lat)

    density = density_bench(solver, sc, cpu=not a.no_cpu_baseline) if not a.graph else None
    solver.close()
    batched = batch_bench(a.batch_scenes, a.batch_steps, 3) if a.batch_scenes > 0 else None
    served = serve_bench(a.serve_seconds, rate=a.serve_rate, plain=a.serve_plain,
                         saturate=a.serve_saturate) if a.serve_seconds > 0 else None
    small = small_scene_bench(cpu=not a.no_cpu_baseline) if a.workload == "apic_slab" and not a.no_small else None

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle on all host cores, then serial on a sub-box) ...")
        cpu = cpu_baseline(sc)
        log(f"cpu baseline {cpu['value']:.3e} p-u/s on {cpu['cores']} threads, serial {cpu['serial']['value']:.3e}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene generator)",
        "config": {"workload": sc.name, "particles": n, "grid": list(sc.desc.resolution),
                   "interp": "quadratic" if sc.desc.interp == 1 else ("cubic" if sc.desc.interp == 2 else "linear"),
                   "transfer": "APIC" if apic else "PIC-FLIP", "dt": sc.desc.dt, "seed": sc.seed,
                   "graph": bool(a.graph), "transfer_fusion": not a.no_fusion, "parallelism": "single GPU"},
        "hbm_roofline_frac_step": step_frac,
        "algorithmic_bytes_per_step": step_bytes,
        "active_nodes_mean": act,
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "kernels": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "density_raster": density,
        "small_scene": small,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
