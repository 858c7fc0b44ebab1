"""GPU sort keys / permutation (bit-exact) and the device 3x3 SVD."""
import numpy as np
import pytest

from mpm_amd import native as N
from mpm_amd import scenes
from mpm_amd.solver import Solver, debug_svd3
from oracle import pyoracle
from tests.svd_cases import special_matrices

pytestmark = pytest.mark.gpu


def _check_sort(gpu, desc, tag):
    gpu.sort()
    keys, ids, seq = gpu.read_sort()
    pos = gpu.read_particles().position
    k_seq = pyoracle.keys(desc, pos[seq])             # keys in the sort's input order
    order = pyoracle.stable_argsort(k_seq)            # std::stable_sort
    ref_ids = seq[order]
    ref_keys = k_seq[order]
    n_esc = int((keys == gpu.host_keys(np.array([[np.nan] * 3], np.float32))[0]).sum())
    print(f"[{tag}] N={keys.size} distinct keys={np.unique(keys).size} escaped={n_esc}", flush=True)
    assert np.array_equal(keys, ref_keys)
    assert np.array_equal(ids, ref_ids)
    assert np.array_equal(np.sort(ids), np.arange(keys.size, dtype=np.uint32))


@pytest.mark.parametrize("interp", [N.LINEAR, N.QUADRATIC_BSPLINE, N.CUBIC_BSPLINE])
def test_sort_bit_exact_after_steps(interp):
    sc = scenes.snow_cube(seed=21, n=20000, res=48, interp=interp, velocity=(1.0, -3.0, 0.5), dt=5e-4)
    gpu = Solver(sc.desc, sc.particles)
    _check_sort(gpu, sc.desc, f"sort interp={interp} t=0")
    gpu.step(5)
    _check_sort(gpu, sc.desc, f"sort interp={interp} t=5")


def test_sort_escapees_and_boundaries():
    # particles on exact cell boundaries, outside the grid, and NaN
    rng = scenes.Rng(5)
    res = 16
    h = 1.0 / res
    lat = (np.indices((8, 8, 8)).reshape(3, -1).T * 2 * h + 0.5 * h).astype(np.float32)  # x/h - 0.5 integer
    out = rng.uniform(300, -0.5, 1.5).reshape(100, 3).astype(np.float32)                   # some outside
    edge = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [-h, 0.5, 0.5], [1 + h, 0.5, 0.5],
                     [np.nan, 0.5, 0.5], [0.5, np.inf, 0.5]], np.float32)
    pos = np.concatenate([lat, out, edge])
    sc = scenes.snow_cube(seed=0, n=pos.shape[0], res=res)
    p = sc.particles
    p.position[:] = pos
    gpu = Solver(sc.desc, p)
    _check_sort(gpu, sc.desc, "sort edges")
    assert gpu.stats()["escaped"] > 0
    gpu.step(2)  # must run without faults with escaped / NaN particles
    st = gpu.read_particles()
    assert np.isfinite(st.position[: lat.shape[0]]).all()


@pytest.mark.parametrize("res,n", [(128, 300_001), (256, 300_001), (128, 1_100_003), (256, 1_100_003)])
def test_sort_large_random(res, n):
    # below 2^20 keys: 128^3 has 22-bit keys, three 8/7/7-bit passes; 256^3
    # 25-bit keys, three 9/8/8-bit passes (512-bin histograms).  From 2^20
    # keys digits are at most 7 bits: 6/6/5/5 and 7/6/6/6.
    sc = scenes.snow_cube(seed=1, n=n, res=res, lo=0.0, hi=1.0)
    gpu = Solver(sc.desc, sc.particles)
    _check_sort(gpu, sc.desc, f"sort {n} random res={res}")


def _svd_invariants(A, U, s, V, tol):
    A = A.astype(np.float64)
    U = U.astype(np.float64)
    V = V.astype(np.float64)
    s = s.astype(np.float64)
    rec = U @ (s[:, :, None] * np.swapaxes(V, 1, 2))
    scale = np.maximum(np.abs(A).reshape(-1, 9).max(1), 1e-30)
    rec_err = np.abs(rec - A).reshape(-1, 9).max(1) / scale
    orthU = np.abs(U @ np.swapaxes(U, 1, 2) - np.eye(3)).reshape(-1, 9).max(1)
    orthV = np.abs(V @ np.swapaxes(V, 1, 2) - np.eye(3)).reshape(-1, 9).max(1)
    assert rec_err.max() <= tol, rec_err.max()
    assert orthU.max() <= 1e-5 and orthV.max() <= 1e-5, (orthU.max(), orthV.max())
    assert (np.linalg.det(U) > 0).all() and (np.linalg.det(V) > 0).all()
    assert (s[:, 0] >= s[:, 1] - 1e-6 * scale).all() and (s[:, 1] >= np.abs(s[:, 2]) - 1e-6 * scale).all()
    return rec_err


def test_device_svd_special_and_random():
    specials = special_matrices()
    rng = scenes.Rng(123)
    rnd = rng.uniform(9 * 4000, -1, 1).reshape(-1, 3, 3)
    near_id = np.eye(3)[None] + rng.uniform(9 * 4000, -0.03, 0.03).reshape(-1, 3, 3)
    A = np.concatenate([specials, rnd, near_id]).astype(np.float32)
    U, s, V = debug_svd3(A)
    rec = _svd_invariants(A, U, s, V, 2e-6)
    # singular values against the FP64 oracle SVD (LAPACK-checked on the CPU)
    _, s64, _ = pyoracle.svd3(A.astype(np.float64), 64)
    sv_err = np.abs(s.astype(np.float64) - s64).max(1) / np.maximum(np.abs(A).reshape(-1, 9).max(1), 1e-30)
    print(f"[svd] n={A.shape[0]} max reconstruction {rec.max():.2e} max singular-value err {sv_err.max():.2e}")
    assert sv_err.max() <= 2e-6
    # sign convention: sign(s3) = sign(det A) where det is not ~0
    d = np.linalg.det(A.astype(np.float64))
    big = np.abs(d) > 1e-3
    assert (np.sign(s[big, 2]) == np.sign(d[big])).all()
    # polar rotation U V^T matches the FP64 one where singular values are distinct
    U64, s64, V64 = pyoracle.svd3(A.astype(np.float64), 64)
    R = U.astype(np.float64) @ np.swapaxes(V.astype(np.float64), 1, 2)
    R64 = U64 @ np.swapaxes(V64, 1, 2)
    gap = np.minimum(np.abs(s64[:, 0] + s64[:, 1]), np.abs(s64[:, 1] + s64[:, 2])) / np.maximum(s64[:, 0], 1e-30)
    ok = gap > 1e-2
    assert np.abs(R[ok] - R64[ok]).max() <= 1e-4
