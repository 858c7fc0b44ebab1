"""Pins of the oracle's density raster (writeToOpenVDB,
src/mpm_solver3d.cu:510-563) and of the density block-file reader.

What fixes each value without running the reference:
- one particle on a node: the B-spline values N(0) = 3/4, N(+-1) = 1/8
  (quadratic), 2/3 and 1/6 (cubic), 1 and 0 (linear);
- a particle half-way between nodes: quadratic N(+-1/2) = 1/2, N(+-3/2) = 0;
- a regular lattice of n^3 particles per cell: every interior node holds
  n^3 m / h^3 (sum_k N(t + k/n) = n per axis);
- partition of unity: sum density h^3 = sum m when every stencil is inside;
- an independent numpy raster written from the B-spline formulas;
- the oracle's own P2G mass grid after initialize() (src/mpm_solver3d.cu:136-167).
"""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd.density import DensityBlocks, read_density_file
from mpm_amd.solver import ParticleInput, SolverDesc
from oracle import pyoracle

H = 1.0 / 64


def _desc(interp, res=32, origin=(0.0, 0.0, 0.0)):
    return SolverDesc(origin=origin, resolution=(res, res, res), stride=H, interp=interp)


def _at(d, x, y, z):
    rx, ry, _ = d.resolution
    return x + rx * (y + ry * z)


@pytest.mark.parametrize("interp,vals,reach", [
    (N.QUADRATIC_BSPLINE, {0: 0.75, 1: 0.125}, 1),
    (N.CUBIC_BSPLINE, {0: 2.0 / 3.0, 1: 1.0 / 6.0, 2: 0.0}, 2),
    (N.LINEAR, {0: 1.0, 1: 0.0}, 1),
])
@pytest.mark.parametrize("precision", [32, 64])
def test_single_particle_on_node(interp, vals, reach, precision):
    d = _desc(interp)
    m = 0.37
    node = (10, 12, 9)
    pos = np.array([[node[0] * H, node[1] * H, node[2] * H]], np.float32)
    dens, act = pyoracle.density(d, pos, m, precision, with_active=True)
    # the stencil range ceil(x/h - r)..floor(x/h + r) on every axis
    assert act.sum() == (2 * reach + 1) ** 3
    for a in range(-reach, reach + 1):
        for b in range(-reach, reach + 1):
            for c in range(-reach, reach + 1):
                want = m * vals[abs(a)] * vals[abs(b)] * vals[abs(c)] / H ** 3
                got = dens[_at(d, node[0] + a, node[1] + b, node[2] + c)]
                assert got == pytest.approx(want, rel=1e-6, abs=1e-9)
                assert act[_at(d, node[0] + a, node[1] + b, node[2] + c)]
    assert dens.sum() == pytest.approx(m / H ** 3, rel=1e-6)


def test_quadratic_half_cell():
    d = _desc(N.QUADRATIC_BSPLINE)
    pos = np.array([[10.5 * H, 7 * H, 20.5 * H]], np.float32)
    dens, act = pyoracle.density(d, pos, 1.0, 64, with_active=True)
    # x: nodes 9..12 reached, weights 0, 1/2, 1/2, 0; y: 6..8 at 1/8, 3/4, 1/8
    assert act.sum() == 4 * 3 * 4
    wx = {9: 0.0, 10: 0.5, 11: 0.5, 12: 0.0}
    wy = {6: 0.125, 7: 0.75, 8: 0.125}
    wz = {19: 0.0, 20: 0.5, 21: 0.5, 22: 0.0}
    for x, a in wx.items():
        for y, b in wy.items():
            for z, c in wz.items():
                assert dens[_at(d, x, y, z)] == pytest.approx(a * b * c / H ** 3, rel=1e-12, abs=1e-9)


@pytest.mark.parametrize("interp", [N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_lattice_density(interp):
    d = _desc(interp)
    n = 2
    m = float(np.float32(1e-3))  # the raster reads float32 masses
    off = (np.arange(n) + 0.5) / n
    cells = np.arange(8, 24)
    ax = (cells[:, None] + off[None, :]).reshape(-1) * H
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    for prec, tol in ((64, 1e-12), (32, 2e-6)):
        dens = pyoracle.density(d, pos, m, prec)
        want = n ** 3 * m / H ** 3
        for node in ((12, 12, 12), (15, 18, 13), (19, 12, 19)):
            assert dens[_at(d, *node)] == pytest.approx(want, rel=tol)


@pytest.mark.parametrize("interp", [N.LINEAR, N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_mass_conservation(interp):
    rng = np.random.default_rng(5)
    d = _desc(interp)
    pos = rng.uniform(4 * H, 28 * H, (500, 3)).astype(np.float32)
    m = rng.uniform(0.5, 2.0, 500).astype(np.float32)
    for prec, tol in ((64, 1e-12), (32, 1e-5)):
        dens = pyoracle.density(d, pos, m, prec)
        assert dens.sum() * H ** 3 == pytest.approx(float(m.astype(np.float64).sum()), rel=tol)


def _bspline(t, interp):
    a = np.abs(t)
    if interp == N.LINEAR:
        return np.where(a < 1, 1 - a, 0.0)
    if interp == N.QUADRATIC_BSPLINE:
        return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    return np.where(a < 1, 0.5 * a ** 3 - a * a + 2.0 / 3.0, np.where(a < 2, (2 - a) ** 3 / 6.0, 0.0))


def _numpy_raster(d, pos, m):
    """Independent raster: the same stencil range (evaluated in float32, as the
    reference's host code does) and B-spline weights in FP64."""
    r = {N.LINEAR: 1.0, N.QUADRATIC_BSPLINE: 1.5, N.CUBIC_BSPLINE: 2.0}[d.interp]
    res = np.array(d.resolution)
    h32 = np.float32(d.stride)
    out = np.zeros(int(np.prod(res)))
    act = np.zeros(int(np.prod(res)), bool)
    for p in range(pos.shape[0]):
        u = pos[p].astype(np.float32) / h32
        lo = np.ceil(u - np.float32(r)).astype(int)
        hi = np.floor(u + np.float32(r)).astype(int)
        axes = []
        for i in range(3):
            idx = np.arange(lo[i], hi[i] + 1)
            idx = idx[(idx >= 0) & (idx < res[i])]
            gp = float(d.origin[i]) + float(d.stride) * idx
            axes.append((idx, _bspline((float(pos[p, i]) - gp) / float(d.stride), d.interp)))
        (ix, wx), (iy, wy), (iz, wz) = axes
        if not (ix.size and iy.size and iz.size):
            continue
        ids = (ix[:, None, None] + res[0] * (iy[None, :, None] + res[1] * iz[None, None, :])).ravel()
        w = (wx[:, None, None] * wy[None, :, None] * wz[None, None, :]).ravel()
        np.add.at(out, ids, float(m[p]) * w / float(d.stride) ** 3)
        act[ids] = True
    return out, act


@pytest.mark.parametrize("interp", [N.LINEAR, N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_matches_independent_numpy_raster(interp):
    rng = np.random.default_rng(11)
    d = _desc(interp, res=16)
    # interior, near-wall and partly-outside particles (stencils clipped)
    pos = rng.uniform(-2 * H, 18 * H, (300, 3)).astype(np.float32)
    m = rng.uniform(0.1, 1.0, 300).astype(np.float32)
    want, want_act = _numpy_raster(d, pos, m)
    got64, act = pyoracle.density(d, pos, m, 64, with_active=True)
    got32 = pyoracle.density(d, pos, m, 32)
    assert np.array_equal(act, want_act)
    scale = np.abs(want).max()
    assert np.abs(got64 - want).max() <= 1e-12 * scale
    assert np.abs(got32 - want).max() <= 1e-5 * scale


@pytest.mark.parametrize("interp", [N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_equals_oracle_p2g_mass(interp):
    rng = np.random.default_rng(3)
    d = _desc(interp, res=24)
    pos = rng.uniform(5 * H, 19 * H, (800, 3)).astype(np.float32)
    parts = ParticleInput(pos, np.zeros_like(pos), 2e-4, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    o = pyoracle.OracleSolver(d, parts, 64)
    mass = o.read_grid()["mass"]  # P2GMass of initialize(), read back as float32
    dens = pyoracle.density(d, pos, parts.mass, 64)
    assert np.abs(dens * H ** 3 - mass).max() <= 1e-6 * mass.max()


def test_block_file_reader(tmp_path):
    # a hand-made file in the documented layout: two blocks on a 6^3 grid
    first = np.array([[0, 0, 0], [4, 0, 4]], np.int32)
    vals = np.arange(128, dtype=np.float32).reshape(2, 4, 4, 4)
    p = tmp_path / "d.mpmd"
    with open(p, "wb") as f:
        f.write(b"MPMDENS1")
        f.write(np.array([6, 6, 6], "<i4").tobytes())
        f.write(np.array([0.0, 0.0, 0.0, 0.5], "<f4").tobytes())
        f.write(np.array([4], "<i4").tobytes())
        f.write(np.array([2], "<i8").tobytes())
        for b in range(2):
            f.write(first[b].tobytes())
            f.write(vals[b].tobytes())
    blk = read_density_file(str(p))
    assert blk.resolution == (6, 6, 6) and blk.stride == 0.5 and blk.count == 2
    dense = blk.dense().reshape(6, 6, 6)  # [z, y, x]
    assert dense[0, 0, 0] == 0.0 and dense[1, 2, 3] == vals[0, 1, 2, 3]
    assert dense[4, 1, 5] == vals[1, 0, 1, 1]
    assert dense[5, 3, 4] == vals[1, 1, 3, 0]
    # nodes of the second block beyond res are dropped by dense()
    assert dense.sum() == pytest.approx(vals[0].sum() + vals[1, :2, :, :2].sum())
    with open(p, "r+b") as f:
        f.write(b"XXXXXXXX")
    with pytest.raises(ValueError):
        read_density_file(str(p))


def test_blocks_dense_empty():
    b = DensityBlocks((8, 8, 8), (0, 0, 0), 0.125, np.zeros((0, 3), np.int32), np.zeros((0, 4, 4, 4), np.float32))
    assert b.dense().shape == (512,) and not b.dense().any()
