"""Pins of the CPU oracle: each oracle function is checked against something
the mathematics fixes independently of the oracle's own formulas (closed
forms, identities, LAPACK, finite differences, hand-computed cases)."""
import numpy as np
import pytest

from mpm_amd import native as N
from oracle import pyoracle as O
from tests.svd_cases import rot, special_matrices

FLT_EPS = np.float32(np.finfo(np.float32).eps)


# ------------------------------------------------------------ B-spline weights
@pytest.mark.parametrize("prec", [32, 64])
def test_bspline_closed_forms(prec):
    x = np.array([0.0, 0.5, -0.5, 1.0, -1.0, 1.5, -1.5, 2.0, 0.25])
    q, qd = O.weights(N.QUADRATIC_BSPLINE, x, prec)
    np.testing.assert_allclose(q[:8], [0.75, 0.5, 0.5, 0.125, 0.125, 0.0, 0.0, 0.0], atol=1e-7)
    np.testing.assert_allclose(qd[:5], [0.0, -1.0, 1.0, -0.5, 0.5], atol=1e-7)
    c, cd = O.weights(N.CUBIC_BSPLINE, x, prec)
    np.testing.assert_allclose(c[[0, 3, 4, 7]], [2 / 3, 1 / 6, 1 / 6, 0.0], atol=1e-7)
    np.testing.assert_allclose(cd[[0, 3, 4, 7]], [0.0, -0.5, 0.5, 0.0], atol=1e-7)
    l, ld = O.weights(N.LINEAR, x, prec)
    np.testing.assert_allclose(l[[0, 1, 2, 3, 8]], [1.0, 0.5, 0.5, 0.0, 0.75], atol=1e-7)
    np.testing.assert_allclose(ld[[0, 1, 2, 3, 8]], [0.0, -1.0, 1.0, 0.0, -1.0], atol=1e-7)


@pytest.mark.parametrize("interp,rng,k", [(N.LINEAR, 1, None), (N.QUADRATIC_BSPLINE, 2, 0.25),
                                          (N.CUBIC_BSPLINE, 2, 1 / 3)])
def test_bspline_moments(interp, rng, k):
    """Partition of unity, zero first moment, sum of derivatives zero and the
    APIC second moment k (1/4 quadratic, 1/3 cubic) that fixes D = k h^2 I."""
    u = np.random.default_rng(0).uniform(0, 1, 500)
    off = np.arange(-rng - 1, rng + 2)
    d = (u[:, None] - off[None, :]).ravel()
    w, wd = O.weights(interp, d, 64)
    w = w.reshape(u.size, -1)
    wd = wd.reshape(u.size, -1)
    dist = off[None, :] - u[:, None]
    np.testing.assert_allclose(w.sum(1), 1.0, atol=1e-12)
    np.testing.assert_allclose((w * dist).sum(1), 0.0, atol=1e-12)
    np.testing.assert_allclose(wd.sum(1), 0.0, atol=1e-12)
    if k is not None:
        np.testing.assert_allclose((w * dist * dist).sum(1), k, atol=1e-12)
    w32, _ = O.weights(interp, d, 32)
    np.testing.assert_allclose(w32.reshape(u.size, -1).sum(1), 1.0, atol=2e-6)


@pytest.mark.parametrize("interp", [N.LINEAR, N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_bspline_derivative_is_finite_difference(interp):
    x = np.random.default_rng(1).uniform(-2.2, 2.2, 400)
    brk = np.array([-2, -1.5, -1, -0.5, 0, 0.5, 1, 1.5, 2])
    x = x[np.min(np.abs(x[:, None] - brk[None, :]), 1) > 1e-3]
    eps = 1e-6
    _, d = O.weights(interp, x, 64)
    wp, _ = O.weights(interp, x + eps, 64)
    wm, _ = O.weights(interp, x - eps, 64)
    np.testing.assert_allclose(d, (wp - wm) / (2 * eps), atol=1e-6)


def test_quadratic_fp32_matches_single_rounding():
    """For |x| < 0.5 the reference evaluates 0.75 - fl(x*x) in double and
    rounds once; the FP32 restatement must give exactly that float."""
    x = np.random.default_rng(2).uniform(-0.49, 0.49, 1000).astype(np.float32)
    w32, _ = O.weights(N.QUADRATIC_BSPLINE, x.astype(np.float64), 32)
    expect = (0.75 - (x * x).astype(np.float64)).astype(np.float32)
    assert np.array_equal(w32.astype(np.float32), expect)


# ---------------------------------------------------------------------- SVD
def _polar(A):
    U, s, Vt = np.linalg.svd(A)
    D = np.diag([1.0, 1.0, np.sign(np.linalg.det(U @ Vt)) or 1.0])
    return U @ D @ Vt


def test_svd_against_lapack_and_invariants():
    rng = np.random.default_rng(3)
    A = np.concatenate([special_matrices(), rng.uniform(-1, 1, (500, 3, 3)),
                        np.eye(3) + rng.uniform(-0.03, 0.03, (500, 3, 3))])
    U, s, V = O.svd3(A, 64)
    for i in range(A.shape[0]):
        a = A[i]
        sn = np.linalg.svd(a, compute_uv=False)
        scale = max(np.abs(a).max(), 1e-300)
        np.testing.assert_allclose(np.abs(s[i]), sn, atol=1e-12 * scale + 1e-300)
        np.testing.assert_allclose(U[i] @ np.diag(s[i]) @ V[i].T, a, atol=1e-12 * scale + 1e-300)
        np.testing.assert_allclose(U[i] @ U[i].T, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(V[i] @ V[i].T, np.eye(3), atol=1e-12)
        assert np.linalg.det(U[i]) > 0 and np.linalg.det(V[i]) > 0
        assert s[i, 0] >= s[i, 1] >= abs(s[i, 2]) - 1e-12 * scale
        d = np.linalg.det(a)
        if abs(d) > 1e-9 * scale ** 3:
            assert np.sign(s[i, 2]) == np.sign(d)
            if d > 0 and sn[0] - sn[1] > 1e-6 and sn[1] - sn[2] > 1e-6:
                np.testing.assert_allclose(U[i] @ V[i].T, _polar(a), atol=1e-9)


def test_svd_fp32_rounds_fp64():
    A = np.eye(3)[None] + np.random.default_rng(4).uniform(-0.02, 0.02, (200, 3, 3))
    U, s, V = O.svd3(A, 32)
    U6, s6, V6 = O.svd3(A.astype(np.float32).astype(np.float64), 64)
    np.testing.assert_allclose(s, s6, atol=1e-7)


# ------------------------------------------------------------------- stress
def _fc_energy(F, Fp, mu0, la0, xi):
    Jp = np.linalg.det(Fp)
    mu = mu0 * np.exp(xi * (1 - Jp))
    la = la0 * np.exp(xi * (1 - Jp))
    s = np.linalg.svd(F, compute_uv=False) * np.sign([1, 1, np.linalg.det(F)])
    J = np.linalg.det(F)
    return mu * ((s - 1) ** 2).sum() + 0.5 * la * (J - 1) ** 2


def _nh_energy(F, Fp, mu0, la0, xi):
    Jp = np.linalg.det(Fp)
    mu = mu0 * np.exp(xi * (1 - Jp))
    la = la0 * np.exp(xi * (1 - Jp))
    J = np.linalg.det(F)
    return mu * (0.5 * (np.trace(F.T @ F) - 3) - np.log(J)) + 0.5 * la * np.log(J) ** 2


def _fd_PFt(energy, F, *args):
    P = np.zeros((3, 3))
    e = 1e-6
    for i in range(3):
        for j in range(3):
            Fp_ = F.copy(); Fp_[i, j] += e
            Fm_ = F.copy(); Fm_[i, j] -= e
            P[i, j] = (energy(Fp_, *args) - energy(Fm_, *args)) / (2 * e)
    return P @ F.T


@pytest.mark.parametrize("model,energy", [(N.FIXED_COROTATED, _fc_energy), (N.NEO_HOOKEAN, _nh_energy)])
def test_stress_is_energy_derivative(model, energy):
    """P F^T of the oracle equals the finite-difference derivative of the
    reference's own energy definitions (src/material_point3d.cu:73-107)."""
    rng = np.random.default_rng(5)
    mu0, la0, xi = 58333.33, 38888.89, 10.0
    for _ in range(20):
        F = rot(rng.normal(size=3), rng.uniform(0, 3)) @ np.diag(rng.uniform(0.9, 1.1, 3)) @ \
            rot(rng.normal(size=3), rng.uniform(0, 3)).T
        Fp = np.eye(3) + rng.uniform(-0.01, 0.01, (3, 3))
        got = O.stress(F, Fp, mu0, la0, xi, model, 64)
        want = _fd_PFt(energy, F, Fp, mu0, la0, xi)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4 * mu0 * 1e-3)


def test_stress_closed_forms():
    mu0, la0, xi = 58333.33, 38888.89, 10.0
    R = rot([1, 2, 3], 0.9)
    assert np.abs(O.stress(R, np.eye(3), mu0, la0, xi, N.FIXED_COROTATED, 64)).max() < 1e-8
    s = 1.01
    got = O.stress(s * np.eye(3), np.eye(3), mu0, la0, xi, N.FIXED_COROTATED, 64)
    np.testing.assert_allclose(got, (2 * mu0 * (s - 1) * s + la0 * (s ** 3 - 1) * s ** 3) * np.eye(3), rtol=1e-12)
    a = 0.99  # hardening: Jp = a^3 scales mu, lambda by exp(xi (1 - a^3))
    got_h = O.stress(s * np.eye(3), a * np.eye(3), mu0, la0, xi, N.FIXED_COROTATED, 64)
    np.testing.assert_allclose(got_h, got * np.exp(xi * (1 - a ** 3)), rtol=1e-12)


def test_lame_from_young_poisson():
    from mpm_amd.solver import ParticleInput, SolverDesc
    p = ParticleInput(np.array([[0.5, 0.5, 0.5]]), np.zeros((1, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    o = O.OracleSolver(SolverDesc(resolution=(16, 16, 16), stride=1 / 16), p, 64)
    st = o.read_particles()
    np.testing.assert_allclose(st.mu0[0], 1.4e5 / (2 * 1.2), rtol=1e-7)
    np.testing.assert_allclose(st.lambda0[0], 1.4e5 * 0.2 / (1.2 * 0.6), rtol=1e-7)


# ---------------------------------------------------------------- collision
G0 = np.zeros(3, np.float32)
G1 = np.ones(3, np.float32)


@pytest.mark.parametrize("pos,vel,mu,expect", [
    ((0.5, 0.5, 0.5), (1.0, -2.0, 0.0), 0.9, (1.0, -2.0, 0.0)),     # interior: untouched
    ((0.5, 0.0, 0.5), (1.0, -2.0, 0.0), 0.9, (0.0, 0.0, 0.0)),      # floor, sticks
    ((0.5, 0.0, 0.5), (3.0, -1.0, 0.0), 0.5, (2.5, 0.0, 0.0)),      # floor, slides with friction
    ((0.5, 0.0, 0.5), (0.3, 1.0, 0.0), 0.5, (0.3, 1.0, 0.0)),       # separating
    ((0.0, 0.0, 0.5), (-1.0, -1.0, 0.5), 0.9, (0.0, 0.0, 0.0)),     # corner, unnormalised normal (1,1,0)
    ((0.0, 0.0, 0.5), (-1.0, -1.0, 0.5), 0.1, tuple(np.array([1.0, 1.0, 0.5]) * (1 - 0.2 / 1.5))),
    ((0.99, 0.5, 0.5), (5.0, 0.0, 0.0), 0.9, (5.0, 0.0, 0.0)),      # just inside the upper wall
    ((1.0, 0.5, 0.5), (5.0, 1.0, 0.0), 0.0, (0.0, 1.0, 0.0)),       # at the upper wall, frictionless
])
def test_boundary_collision_cases(pos, vel, mu, expect):
    got = O.collide(np.array(pos), np.array(vel), G0, G1, mu)
    np.testing.assert_allclose(got, expect, atol=1e-6)


# ---------------------------------------------------------- initialize / step
def _lattice(res, ppc_axis, lo_cell, hi_cell):
    h = 1.0 / res
    c = np.arange(lo_cell * ppc_axis, hi_cell * ppc_axis) + 0.5
    g = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3) / ppc_axis * h
    return g


def test_initialize_lattice_volume():
    """Regular lattice with n^3 particles per cell: interior V0 = h^3 / n^3."""
    from mpm_amd.solver import ParticleInput, SolverDesc
    res, n = 16, 2
    pos = _lattice(res, n, 1, 15)  # interior margin >= 4 cells (cubic stencil reach)
    p = ParticleInput(pos, np.zeros_like(pos), 1e-3, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    for interp in (N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE):
        d = SolverDesc(resolution=(res,) * 3, stride=1 / res, interp=interp)
        st = O.OracleSolver(d, p, 64).read_particles()
        inner = np.all((pos > 6.5 / res) & (pos < 9.5 / res), axis=1)
        np.testing.assert_allclose(st.volume0[inner], (1 / res) ** 3 / n ** 3, rtol=1e-6)


@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_single_particle_free_fall(transfer):
    from mpm_amd.solver import ParticleInput, SolverDesc
    d = SolverDesc(resolution=(32, 32, 32), stride=1 / 32, transfer=transfer, dt=1e-3, gravity=9.8,
                   interp=N.QUADRATIC_BSPLINE)
    p = ParticleInput(np.array([[0.5, 0.6, 0.4]]), np.zeros((1, 3)), 1.0, 10.0, 1.4e5, 0.2, 0.025, 0.0075, 0)
    for prec, tol in ((64, 1e-12), (32, 2e-6)):
        o = O.OracleSolver(d, p, prec)
        n = 20
        o.step(n)
        st = o.read_particles()
        np.testing.assert_allclose(st.velocity[0], [0, -9.8 * n * 1e-3, 0], atol=tol * 10)
        np.testing.assert_allclose(st.position[0, 1], 0.6 - 9.8 * 1e-6 * n * (n + 1) / 2, atol=tol)
        np.testing.assert_allclose(st.Fe[0], np.eye(3), atol=tol)


def test_p2g_conserves_mass_and_momentum():
    from mpm_amd import scenes
    for transfer in (N.PIC_FLIP, N.APIC):
        sc = scenes.snow_cube(seed=9, n=2000, res=32, transfer=transfer, velocity=(0.3, -0.2, 0.5))
        o = O.OracleSolver(sc.desc, sc.particles, 64)
        o.substep("reset")
        o.substep("p2g")
        g = o.read_grid()
        st = o.read_particles()
        m = st.mass.astype(np.float64)
        np.testing.assert_allclose(g["mass"].astype(np.float64).sum(), m.sum(), rtol=1e-6)
        np.testing.assert_allclose(g["velocity"].astype(np.float64).sum(0), (m[:, None] * st.velocity).sum(0),
                                   rtol=1e-6, atol=1e-9)


def test_apic_affine_field_is_exact():
    """APIC reproduces an affine velocity field: with v_p = a + A x_p and
    B_p = A D (so C = A), one transfer gives back v_p = a + A x_p and
    B_p = A D (no stress, no gravity, interior particles)."""
    from mpm_amd.solver import ParticleState, SolverDesc
    res = 32
    h = 1 / res
    pos = _lattice(res, 2, 10, 22)
    n = pos.shape[0]
    a = np.array([0.1, -0.2, 0.05])
    A = np.array([[0.0, 0.3, -0.1], [0.2, 0.0, 0.4], [-0.3, 0.1, 0.0]])
    for interp, k in ((N.QUADRATIC_BSPLINE, 0.25), (N.CUBIC_BSPLINE, 1 / 3)):
        D = k * h * h
        st = ParticleState.empty(n)
        st.position[:] = pos
        st.velocity[:] = a + pos @ A.T
        st.mass[:] = 1e-3
        st.volume0[:] = h ** 3 / 8
        st.Fe[:] = np.eye(3)
        st.Fp[:] = np.eye(3)
        st.mu0[:] = 0.0
        st.lambda0[:] = 0.0
        st.hardening[:] = 10
        st.critical_compression[:] = 0.5
        st.critical_stretch[:] = 0.5
        st.B[:] = A * D
        d = SolverDesc(resolution=(res,) * 3, stride=h, transfer=N.APIC, interp=interp, dt=1e-6, gravity=0.0)
        o = O.OracleSolver(d, st, 64, run_initialize=False)
        o.substep("reset")
        o.substep("p2g")
        o.substep("gravity")
        o.substep("update", 1e-6)
        o.substep("collision", 1e-6)
        o.substep("g2p", 1e-6, 0.0)
        r = o.read_particles64()
        inner = np.all((pos > 13 * h) & (pos < 19 * h), axis=1)
        np.testing.assert_allclose(r["velocity"][inner], (a + pos @ A.T)[inner], atol=1e-7)
        np.testing.assert_allclose(r["B"][inner], np.broadcast_to(A * D, (inner.sum(), 3, 3)), atol=1e-10)


def test_fp32_oracle_tracks_fp64():
    from mpm_amd import scenes
    sc = scenes.snow_cube(seed=0, n=3000, res=32)
    o32 = O.OracleSolver(sc.desc, sc.particles, 32)
    o64 = O.OracleSolver(sc.desc, sc.particles, 64)
    o32.step(10)
    o64.step(10)
    a = o32.read_particles()
    b = o64.read_particles64()
    assert np.abs(a.position - b["position"]).max() < 1e-6
    assert np.abs(a.Fe - b["Fe"]).max() < 1e-5


# ---------------------------------------------------------------- keys/sort
def _keys_numpy(desc, pos):
    """Independent restatement of the cell key in numpy float32."""
    shift = {N.LINEAR: 0.0, N.QUADRATIC_BSPLINE: 0.5, N.CUBIC_BSPLINE: 1.0}[desc.interp]
    width = {N.LINEAR: 2, N.QUADRATIC_BSPLINE: 3, N.CUBIC_BSPLINE: 4}[desc.interp]
    res = np.array(desc.resolution)
    kbn = ((res + 3) >> 2) + 1
    with np.errstate(invalid="ignore"):
        inv_h = np.float32(1.0) / np.float32(desc.stride)
        p32 = pos.astype(np.float32)
        if desc.interp == N.LINEAR:
            fb = np.floor(p32 / np.float32(desc.stride))
            # x < h * b exactly (float64 holds h * b exactly; the sign of
            # the difference is exact)
            below = p32.astype(np.float64) - np.float64(np.float32(desc.stride)) * fb.astype(np.float64) < 0
            fb = np.where(below, fb - 1, fb)
        else:
            fb = np.floor(p32 * inv_h - np.float32(shift))
        ok = np.all((fb >= -(width - 1)) & (fb <= res - 1), axis=1)
    b = np.where(ok[:, None], fb, 0).astype(np.int64)
    blk = ((b + 4) >> 2)
    lin = blk[:, 0] + kbn[0] * (blk[:, 1] + kbn[1] * blk[:, 2])
    cell = ((b + 4) & 3)
    key = lin * 64 + cell[:, 0] + 4 * cell[:, 1] + 16 * cell[:, 2]
    esc = int(kbn.prod()) * 64
    return np.where(ok, key, esc).astype(np.uint32)


@pytest.mark.parametrize("stride", [1 / 32, 1 / 30])
@pytest.mark.parametrize("interp", [N.LINEAR, N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_keys_and_stable_sort(interp, stride):
    from mpm_amd.solver import SolverDesc
    rng = np.random.default_rng(6)
    d = SolverDesc(resolution=(20, 33, 17), stride=stride, interp=interp)
    pos = rng.uniform(-0.2, 1.2, (5000, 3)).astype(np.float32)
    pos[:50] = (np.round(pos[:50] * 64) / 64).astype(np.float32)  # exact cell-boundary values
    pos[50, 0] = np.nan
    k = O.keys(d, pos)
    assert np.array_equal(k, _keys_numpy(d, pos))
    order = O.stable_argsort(k)
    assert np.array_equal(order, np.argsort(k, kind="stable").astype(np.uint32))


@pytest.mark.parametrize("interp,r,shift", [(N.QUADRATIC_BSPLINE, 1.5, 0.5), (N.CUBIC_BSPLINE, 2.0, 1.0),
                                            (N.LINEAR, 1.0, 0.0)])
def test_fixed_stencil_matches_reference_range(interp, r, shift):
    """The device stencil starts at floor(fl(x * fl(1/h)) - s) (linear:
    floor(fl(x/h)), lowered by one where x < h * floor(fl(x/h))); the
    reference loops over ceil(fl(x/h) - r)..floor(fl(x/h) + r)
    (src/mpm_solver3d.cu:251-255).  Where the two node sets differ (x/h - r
    within rounding of an integer) every node in one set and not the other
    has weight < 1e-9 and |N'| < 1e-4 (evaluated at the exact d)."""
    rng = np.random.default_rng(12)
    h = np.float32(1 / 200)  # not a power of two: fl(1/h) is inexact
    inv_h = np.float32(1.0) / h
    k = rng.integers(2, 190, 200000)
    eps = rng.uniform(-4e-5, 4e-5, k.size)
    x = ((k + r + eps) * np.float64(h)).astype(np.float32)  # x/h - r near integers
    lo = np.ceil(x / h - np.float32(r)).astype(np.int64)
    hi = np.floor(x / h + np.float32(r)).astype(np.int64)
    W = {N.LINEAR: 2, N.QUADRATIC_BSPLINE: 3, N.CUBIC_BSPLINE: 4}[interp]
    if interp == N.LINEAR:
        base = np.floor(x / h).astype(np.int64)
        base -= (x.astype(np.float64) - np.float64(h) * base < 0).astype(np.int64)
    else:
        base = np.floor(x * inv_h - np.float32(shift)).astype(np.int64)
    odd = np.nonzero((lo != base) | (hi != base + W - 1))[0]
    if interp != N.LINEAR:
        assert odd.size > 0  # the window was hit
    nodes, xs = [], []
    for i in odd:
        for node in set(range(lo[i], hi[i] + 1)) ^ set(range(base[i], base[i] + W)):
            nodes.append(node)
            xs.append(x[i])
    d = (np.asarray(xs, np.float64) - np.float64(h) * np.asarray(nodes)) / np.float64(h)
    n, dn = O.weights(interp, d)
    assert np.abs(n).max() < 1e-9 and np.abs(dn).max() < 1e-4, (np.abs(n).max(), np.abs(dn).max())


@pytest.mark.parametrize("transfer", [N.PIC_FLIP, N.APIC])
def test_threaded_oracle_matches_serial(transfer):
    """The all-cores timing mode (OpenMP loops, atomic grid sums) computes the
    same step as the serial oracle up to the order of the grid sums; with one
    thread it is the serial oracle bit for bit."""
    from mpm_amd import scenes
    from tests.helpers import particle_errors
    sc = scenes.snow_cube(seed=3, n=6000, res=32, transfer=transfer, velocity=(1.0, -2.0, 0.5), dt=2e-4)
    serial = O.OracleSolver(sc.desc, sc.particles, 32)
    one = O.OracleSolver(sc.desc, sc.particles, 32, threads=1)
    many = O.OracleSolver(sc.desc, sc.particles, 32, threads=4)
    for o in (serial, one, many):
        o.step(5)
    a, b, c = serial.read_particles(), one.read_particles(), many.read_particles()
    for f in ("position", "velocity", "Fe", "Fp", "B"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    e = particle_errors(c, a)
    assert all(v <= 1e-5 for v in e.values()), e
    with pytest.raises(ValueError):
        O.OracleSolver(sc.desc, serial.read_particles(), 32, threads=2)


# ------------------------------------------------ deformation-gradient update
# updateDeformationGradient (reference src/material_point3d.cu:53-70): trial
# Fe* = (I + dt grad v) Fe, singular values of Fe* clamped to
# [1 - theta_c, 1 + theta_s], and Fp absorbs the rest so that the product
# Fe Fp follows the trial.  Pinned by what any correct reading must satisfy:
# the product invariant, the clamp range with its two distinct bounds, the
# no-op inside the range and the shared singular vectors.
TH_C, TH_S = 2.5e-2, 7.5e-3


def _rand_rot(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def _with_sv(rng, sv):
    """R1 diag(sv) R2 for random rotations (one per row of sv)."""
    sv = np.asarray(sv, np.float64)
    n = sv.shape[0]
    return _rand_rot(rng, n) @ (sv[:, :, None] * _rand_rot(rng, n))


def test_update_F_product_invariant_and_clamp_range():
    rng = np.random.default_rng(11)
    n = 400
    Fe = _with_sv(rng, rng.uniform(0.95, 1.03, (n, 3)))
    Fp = np.eye(3) + rng.uniform(-0.05, 0.05, (n, 3, 3))
    grad = rng.uniform(-40.0, 40.0, (n, 3, 3))
    dt = 2.5e-4
    trial = (np.eye(3) + dt * grad) @ Fe
    Fe1, Fp1 = O.update_F(Fe, Fp, grad, dt, TH_C, TH_S, 64)
    # Fe Fp follows the trial exactly (real arithmetic; FP64 rounding here)
    prod = Fe1 @ Fp1
    want = trial @ Fp
    assert np.abs(prod - want).max() <= 1e-12 * np.abs(want).max()
    # clamped singular values inside [1 - theta_c, 1 + theta_s]
    s1 = np.linalg.svd(Fe1, compute_uv=False)
    assert s1.max() <= 1 + TH_S + 1e-12 and s1.min() >= 1 - TH_C - 1e-12
    # Fe keeps the trial's singular vectors: U^T Fe V is diagonal with the
    # clamped trial singular values (numpy's SVD of the trial)
    U, s, Vt = np.linalg.svd(trial)
    D = np.swapaxes(U, 1, 2) @ Fe1 @ np.swapaxes(Vt, 1, 2)
    off = D - D * np.eye(3)
    assert np.abs(off).max() <= 1e-12
    np.testing.assert_allclose(np.diagonal(D, axis1=1, axis2=2), np.clip(s, 1 - TH_C, 1 + TH_S), atol=1e-12)
    # the FP32 restatement follows the FP64 one
    Fe32, Fp32 = O.update_F(Fe, Fp, grad, dt, TH_C, TH_S, 32)
    assert np.abs(Fe32 - Fe1).max() < 2e-6 and np.abs(Fp32 - Fp1).max() < 2e-6


def test_update_F_inside_range_is_noop():
    rng = np.random.default_rng(12)
    n = 200
    # trials strictly inside the range (grad = 0: the trial is Fe itself)
    Fe = _with_sv(rng, rng.uniform(1 - TH_C + 1e-3, 1 + TH_S - 1e-3, (n, 3)))
    Fp = np.eye(3) + rng.uniform(-0.05, 0.05, (n, 3, 3))
    Fe1, Fp1 = O.update_F(Fe, Fp, np.zeros((n, 3, 3)), 1e-4, TH_C, TH_S, 64)
    assert np.abs(Fe1 - Fe).max() <= 1e-13
    assert np.abs(Fp1 - Fp).max() <= 1e-13


@pytest.mark.parametrize("sv,expect", [
    ((1.05, 1.0, 0.99), (1 + TH_S, 1.0, 0.99)),          # stretched past 1 + theta_s
    ((1.0, 0.95, 0.90), (1.0, 1 - TH_C, 1 - TH_C)),      # compressed past 1 - theta_c
    ((1.2, 1.0, 0.5), (1 + TH_S, 1.0, 1 - TH_C)),        # both bounds at once
])
def test_update_F_clamps_to_the_named_bound(sv, expect):
    """The critical stretch bounds from above and the critical compression
    from below: swapping them (or clamping to the wrong side) moves the
    end singular values off these exact numbers."""
    rng = np.random.default_rng(13)
    Fe = _with_sv(rng, np.tile(sv, (50, 1)))
    Fp = np.tile(np.eye(3), (50, 1, 1))
    Fe1, Fp1 = O.update_F(Fe, Fp, np.zeros((50, 3, 3)), 1e-4, TH_C, TH_S, 64)
    s1 = np.linalg.svd(Fe1, compute_uv=False)
    np.testing.assert_allclose(s1, np.tile(expect, (50, 1)), atol=1e-12)
    # and the plastic part carries the excess: sv(Fp) = sv(trial) / sv(Fe)
    sp = np.sort(np.linalg.svd(Fp1, compute_uv=False), axis=1)
    np.testing.assert_allclose(sp, np.sort(np.tile(np.array(sv) / np.array(expect), (50, 1)), axis=1), atol=1e-12)


# ------------------------------------------------------- PIC-FLIP blend
def _two_lattices(res, lo, hi, oA, oB, mA, mB, uA, uB):
    """One particle of kind A and one of kind B per cell of [lo, hi)^3, at
    the fixed cell offsets oA, oB; zero Young's modulus (no stress)."""
    from mpm_amd.solver import ParticleInput
    h = 1.0 / res
    c = np.arange(lo, hi)
    cells = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = np.concatenate([(cells + oA) * h, (cells + oB) * h])
    n = cells.shape[0]
    vel = np.concatenate([np.tile(uA, (n, 1)), np.tile(uB, (n, 1))])
    mass = np.concatenate([np.full(n, mA), np.full(n, mB)])
    kind = np.concatenate([np.zeros(n, int), np.ones(n, int)])
    p = ParticleInput(pos, vel, mass, 10.0, 0.0, 0.2, TH_C, TH_S, 0)
    return p, kind, cells


# h = 1/32 and offsets in eighths of a cell: every position is exact in FP32,
# so both lattices are exactly regular
BLEND_SCENE = dict(res=32, lo=6, hi=22, oA=np.array([0.375, 0.625, 0.25]), oB=np.array([0.75, 0.125, 0.5]),
                   mA=2e-3, mB=1e-3, uA=np.array([1.0, 0.0, 0.2]), uB=np.array([-1.0, 0.5, 0.4]))


def blend_expectation(alpha, g, dt):
    """Closed form of one PIC-FLIP step of the two-lattice scene for the
    particles whose whole stencil lies on interior nodes.  Every interior
    node holds mass mA + mB and momentum mA uA + mB uB (each lattice's
    weights sum to 1 at every interior node), so v = (mA uA + mB uB) / (mA +
    mB) and v* = v - g dt y; G2P then gives v_pic = v* and v_flip = u - g dt y
    (partition of unity over the stencil), blended as
    alpha v_flip + (1 - alpha) v_pic."""
    s = BLEND_SCENE
    alpha = float(np.float32(alpha))  # the blend coefficient is a float (include/mpm_solver3d.h:31)
    v = (s["mA"] * s["uA"] + s["mB"] * s["uB"]) / (s["mA"] + s["mB"])
    gy = np.array([0.0, g * dt, 0.0])
    eA = alpha * (s["uA"] - gy) + (1 - alpha) * (v - gy)
    eB = alpha * (s["uB"] - gy) + (1 - alpha) * (v - gy)
    return v, eA, eB


def blend_scene(alpha, g=9.8, dt=1e-4):
    from mpm_amd.solver import SolverDesc
    s = BLEND_SCENE
    p, kind, cells = _two_lattices(s["res"], s["lo"], s["hi"], s["oA"], s["oB"], s["mA"], s["mB"], s["uA"],
                                   s["uB"])
    d = SolverDesc(resolution=(s["res"],) * 3, stride=1.0 / s["res"], transfer=N.PIC_FLIP, blend=alpha,
                   interp=N.QUADRATIC_BSPLINE, dt=dt, gravity=g)
    cc = np.concatenate([cells, cells])
    inner = np.all((cc >= s["lo"] + 3) & (cc < s["hi"] - 3), axis=1)
    return d, p, kind, inner


@pytest.mark.parametrize("alpha", [0.0, 0.95, 1.0])
def test_pic_flip_blend_two_lattices(alpha):
    g, dt = 9.8, 1e-4
    d, p, kind, inner = blend_scene(alpha, g, dt)
    v, eA, eB = blend_expectation(alpha, g, dt)
    for prec, tol in ((64, 1e-12), (32, 2e-6)):
        o = O.OracleSolver(d, p, prec)
        # the grid the closed form describes
        o.substep("reset")
        o.substep("p2g")
        o.substep("gravity")
        o.substep("update", dt)
        o.substep("collision", dt)
        gr = o.read_grid()
        res, s = d.resolution[0], BLEND_SCENE
        idx = np.arange(res)
        X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
        nid = (X + res * (Y + res * Z)).ravel()
        core = ((X >= s["lo"] + 2) & (X < s["hi"] - 2) & (Y >= s["lo"] + 2) & (Y < s["hi"] - 2) &
                (Z >= s["lo"] + 2) & (Z < s["hi"] - 2)).ravel()
        np.testing.assert_allclose(gr["mass"][nid[core]], s["mA"] + s["mB"], rtol=max(tol, 1e-7))
        np.testing.assert_allclose(gr["velocity"][nid[core]], np.tile(v, (core.sum(), 1)), atol=max(tol, 1e-6))
        np.testing.assert_allclose(gr["velocity_star"][nid[core]], np.tile(v - [0, g * dt, 0], (core.sum(), 1)),
                                   atol=max(tol, 1e-6))
        # one full step from the same start
        o = O.OracleSolver(d, p, prec)
        o.step(1)
        vel = o.read_particles64()["velocity"] if prec == 64 else o.read_particles().velocity.astype(np.float64)
        a = inner & (kind == 0)
        b = inner & (kind == 1)
        assert a.sum() > 100 and b.sum() > 100
        np.testing.assert_allclose(vel[a], np.tile(eA, (a.sum(), 1)), atol=tol)
        np.testing.assert_allclose(vel[b], np.tile(eB, (b.sum(), 1)), atol=tol)


# ------------------------------------------------------ light-node branch
# CollocatedGridData3D::updateVelocity (reference src/grid3d.cu:36-41) and
# computeGravityForces (src/mpm_solver3d.cu:394-403) act only on nodes with
# mass > FLT_EPSILON: a lighter node keeps its raw momentum in "velocity",
# its velocity_star stays 0 and it gets no gravity.
def light_node_scene(transfer):
    from mpm_amd.solver import ParticleInput, SolverDesc
    res = 32
    h = 1.0 / res
    x = (np.array([12, 15, 17]) + 0.25) * h   # x/h - i = 0.25 on every axis
    vp = np.array([0.375, -0.25, 0.125])  # exact in FP32
    p = ParticleInput(x[None], vp[None], 1e-6, 10.0, 1.4e5, 0.2, TH_C, TH_S, 0)
    d = SolverDesc(resolution=(res,) * 3, stride=h, transfer=transfer, blend=0.0, interp=N.QUADRATIC_BSPLINE,
                   dt=1e-4, gravity=9.8)
    # quadratic weights at distances 1.25, 0.25, 0.75 (closed forms):
    # N(1.25) = (1.5 - 1.25)^2 / 2 = 1/32, N(0.25) = 3/4 - 1/16, N(0.75) = (1.5 - 0.75)^2 / 2
    w1 = np.array([1 / 32, 0.75 - 0.0625, 0.5 * 0.75 ** 2])
    W = w1[:, None, None] * w1[None, :, None] * w1[None, None, :]
    base = np.array([11, 14, 16])
    return d, p, vp, W, base


def test_light_node_keeps_momentum_no_gravity():
    d, p, vp, W, base = light_node_scene(N.APIC)
    m = 1e-6
    g, dt = 9.8, 1e-4
    heavy = np.float32(m) * W.astype(np.float32) > FLT_EPS
    assert 0 < heavy.sum() < 27  # both branches are exercised
    for prec in (64, 32):
        o = O.OracleSolver(d, p, prec)
        o.substep("reset")
        o.substep("p2g")
        o.substep("gravity")
        o.substep("update", dt)
        o.substep("collision", dt)
        gr = o.read_grid()
        res = d.resolution[0]
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    i = (base[0] + a) + res * ((base[1] + b) + res * (base[2] + c))
                    w = W[a, b, c]
                    np.testing.assert_allclose(gr["mass"][i], m * w, rtol=1e-6)
                    if heavy[a, b, c]:
                        np.testing.assert_allclose(gr["velocity"][i], vp, rtol=1e-5)
                        np.testing.assert_allclose(gr["velocity_star"][i], vp - [0, g * dt, 0], rtol=1e-5)
                        np.testing.assert_allclose(gr["force"][i][1], -g * m * w, rtol=1e-5)
                    else:
                        np.testing.assert_allclose(gr["velocity"][i], m * w * vp, rtol=1e-5)
                        assert np.all(gr["velocity_star"][i] == 0.0)
                        assert gr["force"][i][1] == 0.0


@pytest.mark.parametrize("transfer", [N.APIC, N.PIC_FLIP])
def test_light_node_branch_in_one_step(transfer):
    """After one step (PIC-FLIP with blend 0 or APIC), the particle's velocity
    is the heavy nodes' weight sum times v_p - g dt y: light nodes add 0."""
    d, p, vp, W, base = light_node_scene(transfer)
    g, dt = 9.8, 1e-4
    heavy = np.float32(1e-6) * W.astype(np.float32) > FLT_EPS
    expect = W[heavy].sum() * (vp - [0, g * dt, 0])
    for prec, tol in ((64, 1e-12), (32, 1e-6)):
        o = O.OracleSolver(d, p, prec)
        o.step(1)
        v = o.read_particles64()["velocity"][0] if prec == 64 else o.read_particles().velocity[0].astype(np.float64)
        np.testing.assert_allclose(v, expect, atol=tol * 10, rtol=tol)
