"""Device density raster (replacement of writeToOpenVDB,
src/mpm_solver3d.cu:510-572) against the oracle's restatement of the
reference's host raster, on the same positions.

After a fused step (APIC, quadratic) the raster reads the grid's mass
plane, which the fused kernel has already summed for the next step; the
other cases run the raster's own sort and mass-only P2G.  The same bar
holds for both.

Bar: normwise relative error <= 1e-5 (float atomics reorder the FP32 sums);
the block copy, the dense copy and the file hold the same bits; every node
the oracle gives a nonzero density lies in a returned block.
"""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.batch import Batch
from mpm_amd.density import read_density_file
from mpm_amd.solver import ParticleInput, Solver, SolverDesc
from oracle import pyoracle
from tests.helpers import particle_errors, rel_err, report

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _check_blocks(s: Solver, dense: np.ndarray, ref: np.ndarray):
    blk = s.density_blocks()
    # the block copy reads the same plane as the dense copy
    assert np.array_equal(blk.dense(), dense)
    rx, ry, rz = s.desc.resolution
    covered = np.zeros((rz, ry, rx), bool)
    for f in blk.first:
        covered[f[2]:f[2] + 4, f[1]:f[1] + 4, f[0]:f[0] + 4] = True
    assert not np.any((ref.reshape(rz, ry, rx) != 0) & ~covered)
    assert np.all(blk.first % 4 == 0)
    return blk


@pytest.mark.parametrize("interp,transfer", [(N.QUADRATIC_BSPLINE, N.APIC), (N.QUADRATIC_BSPLINE, N.PIC_FLIP),
                                             (N.CUBIC_BSPLINE, N.APIC), (N.LINEAR, N.PIC_FLIP)])
def test_density_matches_oracle_after_steps(interp, transfer, tmp_path):
    sc = scenes.snow_cube(seed=21, n=6000, res=32, transfer=transfer, interp=interp, velocity=(0.3, -0.2, 0.1))
    s = Solver(sc.desc, sc.particles)
    s.step(3)
    st = s.read_particles()
    print("density case", interp, transfer, flush=True)
    dense = s.density()
    ref = pyoracle.density(sc.desc, st.position, st.mass, 32)
    e = {"density": rel_err(dense, ref), "mass": abs(dense.sum() - ref.sum()) / ref.sum()}
    report(f"density interp={interp} transfer={transfer}", e)
    assert e["density"] <= TOL and e["mass"] <= TOL
    blk = _check_blocks(s, dense, ref)
    p = tmp_path / "frame.mpmd"
    s.write_density(str(p))
    f = read_density_file(str(p))
    assert f.resolution == tuple(sc.desc.resolution) and f.stride == np.float32(sc.desc.stride)
    assert np.array_equal(f.first, blk.first) and np.array_equal(f.values, blk.values)
    assert np.array_equal(f.dense(), dense)


def test_density_clipped_stencils_and_escapes():
    # particles near and beyond every wall: clipped stencils, escaped particles
    rng = np.random.default_rng(8)
    res, h = 24, 1.0 / 24
    pos = rng.uniform(-3 * h, 27 * h, (5000, 3)).astype(np.float32)
    d = SolverDesc(resolution=(res, res, res), stride=h, transfer=N.APIC)
    p = ParticleInput(pos, np.zeros_like(pos), rng.uniform(1e-4, 3e-4, 5000), 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    s = Solver(d, p)
    dense = s.density()
    ref = pyoracle.density(d, p.position, p.mass, 32)
    e = {"density": rel_err(dense, ref)}
    report("density near walls", e)
    assert e["density"] <= TOL
    _check_blocks(s, dense, ref)


def test_density_after_fused_step_equals_standalone_raster():
    sc = scenes.snow_cube(seed=24, n=8000, res=32, transfer=N.APIC, velocity=(0.4, -0.6, 0.2))
    a = Solver(sc.desc, sc.particles)  # fused steps: the raster is the grid's mass plane
    a.step(4)
    da = a.density()
    b = Solver(sc.desc, sc.particles)
    b.set_transfer_fusion(False)  # two-kernel steps: the raster's own P2G
    b.step(4)
    db = b.density()
    e = {"density": rel_err(da, db)}
    report("raster after fused steps vs standalone raster", e)
    assert e["density"] <= TOL
    assert a.density_blocks().count == b.density_blocks().count


def test_density_cached_and_refreshed():
    sc = scenes.snow_cube(seed=22, n=4000, res=32, transfer=N.APIC, velocity=(0.0, -1.0, 0.0))
    s = Solver(sc.desc, sc.particles)
    a = s.density()
    b = s.density()
    assert np.array_equal(a, b)  # no change in between: the same raster
    s.step(5)
    c = s.density()
    ref = pyoracle.density(sc.desc, s.read_particles().position, sc.particles.mass, 32)
    # the second raster starts from a cleared plane
    assert rel_err(c, ref) <= TOL
    assert not np.array_equal(a, c)


def test_density_does_not_disturb_steps():
    sc = scenes.snow_cube(seed=23, n=5000, res=32, transfer=N.APIC, velocity=(0.2, 0.1, -0.3))
    a = Solver(sc.desc, sc.particles)
    b = Solver(sc.desc, sc.particles)
    a.step(2)
    a.density()
    a.step(2)
    b.step(4)
    e = particle_errors(a.read_particles(), b.read_particles())
    report("steps around a raster", e)
    assert all(v <= TOL for v in e.values()), e


def test_density_of_batch_member():
    sc = scenes.snow_cube(seed=24, n=5000, res=32, transfer=N.APIC, velocity=(0.1, -0.2, 0.3))
    other = scenes.snow_cube(seed=25, n=3000, res=32, transfer=N.APIC)
    alone = Solver(sc.desc, sc.particles)
    member = Solver(sc.desc, sc.particles)
    second = Solver(other.desc, other.particles)
    b = Batch(N.QUADRATIC_BSPLINE, N.APIC)
    b.add(member)
    b.add(second)
    b.step(3)
    alone.step(3)
    da, dm = alone.density(), member.density()
    assert rel_err(dm, da) <= TOL
    # the batch keeps stepping correctly after a member's raster
    b.step(2)
    alone.step(2)
    b.synchronize()
    e = particle_errors(member.read_particles(), alone.read_particles())
    report("batch member after a raster", e)
    assert all(v <= TOL for v in e.values()), e


def test_density_empty_scene(tmp_path):
    d = SolverDesc(resolution=(16, 16, 16), stride=1 / 16, transfer=N.APIC)
    empty = ParticleInput(np.zeros((0, 3)), np.zeros((0, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    s = Solver(d, empty)
    assert not s.density().any()
    assert s.density_blocks().count == 0
    s.write_density(str(tmp_path / "e.mpmd"))
    assert read_density_file(str(tmp_path / "e.mpmd")).count == 0


def test_density_conserves_mass_large():
    # 2M particles of the slab workload (all stencils inside): sum rho h^3 = sum m
    sc = scenes.apic_slab(nx=128, ny=128, nz=128, res=256)
    s = Solver(sc.desc, sc.particles)
    s.step(2)
    dense = s.density()
    total = float(sc.particles.mass.astype(np.float64).sum())
    h3 = float(sc.desc.stride) ** 3
    err = abs(float(dense.astype(np.float64).sum()) * h3 - total) / total
    print(f"[density 2M] mass rel err {err:.3e}, blocks {s.density_blocks().count}", flush=True)
    assert err <= TOL
