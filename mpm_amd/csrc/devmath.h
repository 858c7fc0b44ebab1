// devmath.h -- device math of the MPM solver (gfx950, FP32): 3x3 helpers,
// branch-free 3x3 SVD, B-spline weights, boundary collision, sort key.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

#define MPM_FLT_EPS 1.1920928955078125e-07f

// Device math is also compiled for the host so that its numerics can be
// checked on a machine without a GPU.
#define MPM_HD __host__ __device__ __forceinline__

// Reciprocal square root refined by one Newton step (hardware v_rsq_f32 is
// accurate to ~1 ulp; the refinement keeps Givens pairs orthonormal to
// ~1 ulp so rotations accumulated over many sweeps stay rotations).
MPM_HD float mpm_rsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    float y = __builtin_amdgcn_rsqf(x);
#else
    float y = 1.0f / sqrtf(x);
#endif
    return y * (1.5f - 0.5f * x * y * y);
}

// Reciprocal to ~1 ulp (hardware v_rcp_f32) where correct rounding does not
// matter: Jacobi tangents, clamped singular values.
MPM_HD float mpm_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// ------------------------------------------------------------ 3x3 helpers
// Matrices are float[3][3] row-major, fully unrolled so they stay in VGPRs.
MPM_HD void m3_mul(const float a[3][3], const float b[3][3], float c[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) c[i][j] = (a[i][0] * b[0][j] + a[i][1] * b[1][j]) + a[i][2] * b[2][j];
}
// c = a * b^T
MPM_HD void m3_mul_bt(const float a[3][3], const float b[3][3], float c[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) c[i][j] = (a[i][0] * b[j][0] + a[i][1] * b[j][1]) + a[i][2] * b[j][2];
}
MPM_HD float m3_det(const float m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[1][0] * (m[0][1] * m[2][2] - m[0][2] * m[2][1]) +
           m[2][0] * (m[0][1] * m[1][2] - m[0][2] * m[1][1]);
}

// ---------------------------------------------------------------- 3x3 SVD
// A = U diag(s) V^T, U and V rotations, s0 >= s1 >= |s2|, sign(s2) =
// sign(det A) -- the convention of McAdams et al. 2011 ("Computing the
// Singular Value Decomposition of 3x3 matrices with minimal branching and
// elementary floating point operations", UW-Madison TR1690), which the
// reference's third-party SVD implements.  Written from the method:
//   1. symmetric eigenproblem of A^T A by cyclic Jacobi with exact
//      (Rutishauser) rotations: two sweeps, then two more unless the
//      off-diagonal is already below FP32 resolution;
//   2. B = A V, columns sorted by decreasing norm (swap + negate keeps
//      det V = +1);
//   3. Givens QR of B (the pivots end >= 0), s_i = u_i . b_i.
// U and V each get one Newton-Schulz step X <- X (3I - X^T X) / 2, which
// removes the orthogonality drift that rounding leaves after many
// rotations.  Every decision is a select, so the wave never diverges.

#define MPM_SVD_SWEEPS 4

// Jacobi rotation J(p,q) zeroing S_pq (J_pp = J_qq = c, J_pq = -J_qp = s;
// S <- J^T S J, V <- V J); spp/sqq/spq are S_pp, S_qq, S_pq and srp/srq the
// entries coupling the third index.
MPM_HD void jacobi_pair(float& spp, float& sqq, float& spq, float& srp, float& srq, float V[3][3], int p, int q) {
    const float d = sqq - spp;
    const float den = fabsf(d) + sqrtf(d * d + 4.0f * spq * spq);
    const float t = den > 0.0f ? ((d >= 0.0f ? 2.0f : -2.0f) * spq) * mpm_rcp(den) : 0.0f;  // tan(theta)
    const float c = mpm_rsqrt(1.0f + t * t);
    const float s = t * c;
    spp = spp - t * spq;
    sqq = sqq + t * spq;
    spq = 0.0f;
    const float rp = srp, rq = srq;
    srp = c * rp - s * rq;
    srq = s * rp + c * rq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - s * vkq;
        V[k][q] = s * vkp + c * vkq;
    }
}

MPM_HD void newton_schulz(float X[3][3]) {
    float XtX[3][3], Y[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            XtX[i][j] = (i == j ? 1.5f : 0.0f) - 0.5f * ((X[0][i] * X[0][j] + X[1][i] * X[1][j]) + X[2][i] * X[2][j]);
    m3_mul(X, XtX, Y);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[i][j] = Y[i][j];
}

MPM_HD void svd_swap_neg(float B[3][3], float V[3][3], float& ni, float& nj, int i, int j, bool doit) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float bi = B[k][i], bj = B[k][j];
        B[k][i] = doit ? bj : bi;
        B[k][j] = doit ? -bi : bj;
        float vi = V[k][i], vj = V[k][j];
        V[k][i] = doit ? vj : vi;
        V[k][j] = doit ? -vi : vj;
    }
    float a = ni, b = nj;
    ni = doit ? b : a;
    nj = doit ? a : b;
}

// Givens rotation on rows p,q of W zeroing W[q][col] (the pivot becomes
// |(a,b)| >= 0); U accumulates the transposed rotations (W_orig = U W).
MPM_HD void qr_givens(float W[3][3], float U[3][3], int p, int q, int col) {
    const float a = W[p][col], b = W[q][col];
    const float r2 = a * a + b * b;
    const bool ok = r2 > 1.0e-36f;
    const float ir = mpm_rsqrt(ok ? r2 : 1.0f);
    const float c = ok ? a * ir : 1.0f;
    const float s = ok ? b * ir : 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float bp = W[p][k], bq = W[q][k];
        W[p][k] = c * bp + s * bq;
        W[q][k] = -s * bp + c * bq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float up = U[k][p], uq = U[k][q];
        U[k][p] = c * up + s * uq;
        U[k][q] = -s * up + c * uq;
    }
}

MPM_HD void svd3(const float A[3][3], float U[3][3], float s[3], float V[3][3]) {
    // S = A^T A
    float s00 = (A[0][0] * A[0][0] + A[1][0] * A[1][0]) + A[2][0] * A[2][0];
    float s01 = (A[0][0] * A[0][1] + A[1][0] * A[1][1]) + A[2][0] * A[2][1];
    float s02 = (A[0][0] * A[0][2] + A[1][0] * A[1][2]) + A[2][0] * A[2][2];
    float s11 = (A[0][1] * A[0][1] + A[1][1] * A[1][1]) + A[2][1] * A[2][1];
    float s12 = (A[0][1] * A[0][2] + A[1][1] * A[1][2]) + A[2][1] * A[2][2];
    float s22 = (A[0][2] * A[0][2] + A[1][2] * A[1][2]) + A[2][2] * A[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0f : 0.0f;
#pragma unroll
    for (int sweep = 0; sweep < MPM_SVD_SWEEPS / 2; ++sweep) {
        jacobi_pair(s00, s11, s01, s02, s12, V, 0, 1);  // r = 2
        jacobi_pair(s00, s22, s02, s01, s12, V, 0, 2);  // r = 1 (S_10 = s01, S_12 = s12)
        jacobi_pair(s11, s22, s12, s01, s02, V, 1, 2);  // r = 0 (S_01 = s01, S_02 = s02)
    }
    // Cyclic Jacobi converges quadratically: a deformation gradient near a
    // rotation (|A^T A - I| ~ 1e-2) leaves an off-diagonal of ~1e-8 of the
    // diagonal after two sweeps, far below FP32 resolution.  The remaining
    // sweeps run only when the off-diagonal is still above 1e-7 relative
    // (general matrices), so the result is as accurate as the full count.
    {
        const float off = (s01 * s01 + s02 * s02) + s12 * s12;
        const float dia = (s00 * s00 + s11 * s11) + s22 * s22;
        if (off > 1.0e-14f * dia) {
#pragma unroll
            for (int sweep = MPM_SVD_SWEEPS / 2; sweep < MPM_SVD_SWEEPS; ++sweep) {
                jacobi_pair(s00, s11, s01, s02, s12, V, 0, 1);
                jacobi_pair(s00, s22, s02, s01, s12, V, 0, 2);
                jacobi_pair(s11, s22, s12, s01, s02, V, 1, 2);
            }
        }
    }
    newton_schulz(V);
    float B[3][3];
    m3_mul(A, V, B);
    float n0 = (B[0][0] * B[0][0] + B[1][0] * B[1][0]) + B[2][0] * B[2][0];
    float n1 = (B[0][1] * B[0][1] + B[1][1] * B[1][1]) + B[2][1] * B[2][1];
    float n2 = (B[0][2] * B[0][2] + B[1][2] * B[1][2]) + B[2][2] * B[2][2];
    svd_swap_neg(B, V, n0, n1, 0, 1, n0 < n1);
    svd_swap_neg(B, V, n0, n2, 0, 2, n0 < n2);
    svd_swap_neg(B, V, n1, n2, 1, 2, n1 < n2);
    float W[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            W[i][j] = B[i][j];
            U[i][j] = (i == j) ? 1.0f : 0.0f;
        }
    qr_givens(W, U, 0, 1, 0);
    qr_givens(W, U, 0, 2, 0);
    qr_givens(W, U, 1, 2, 1);
    newton_schulz(U);
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = (U[0][i] * B[0][i] + U[1][i] * B[1][i]) + U[2][i] * B[2][i];
}

// ----------------------------------------------------- B-spline weights
// N(d) and N'(d) of src/interpolation3d.cu:88-142 in FP32.  Contraction is
// off so that the quadratic and linear kernels round exactly as the
// reference's double-promoted expressions do (one rounding of an exact
// value); the cubic kernel differs from the reference by at most a few ulp.
template <int INTERP>
MPM_HD void bspline(float d, float& N, float& dN) {
#pragma clang fp contract(off)
    float xa = fabsf(d);
    if (INTERP == MPM_INTERP_LINEAR) {
        N = xa < 1.0f ? 1.0f - xa : 0.0f;
        dN = (xa < 1.0f && xa > MPM_FLT_EPS) ? -copysignf(1.0f, d) : 0.0f;  // -d/|d|
    } else if (INTERP == MPM_INTERP_QUADRATIC_BSPLINE) {
        float t = 1.5f - xa;
        float n_in = 0.75f - xa * xa;
        float n_out = 0.5f * (t * t);
        N = xa < 0.5f ? n_in : (xa < 1.5f ? n_out : 0.0f);
        float g_in = -2.0f * d;
        float g_out = d - copysignf(1.5f, d);
        dN = xa < 0.5f ? g_in : (xa < 1.5f ? g_out : 0.0f);
    } else {
        float x2 = xa * xa;
        float t = 2.0f - xa;
        float n_in = (0.5f * x2 * xa - x2) + (2.0f / 3.0f);
        float n_out = (t * t * t) / 6.0f;
        N = xa < 1.0f ? n_in : (xa < 2.0f ? n_out : 0.0f);
        float g_in = 1.5f * d * xa - 2.0f * d;
        float g_out = (2.0f * d - copysignf(2.0f, d)) - 0.5f * (d * xa);
        dN = xa < 1.0f ? g_in : (xa < 2.0f ? g_out : 0.0f);
    }
}

// The same N(d), N'(d) for node j of the fixed stencil, whose offset d lies
// in a known piece: quadratic j = 0, 1, 2 in [0.5,1.5), [-0.5,0.5),
// [-1.5,-0.5); cubic j = 0..3 in [1,2), [0,1), [-1,0), [-2,-1).  Each piece
// is bspline's expression for that piece, so in-range values are the same
// bits; the piece is chosen by j instead of by comparisons (no branches).
// An offset that rounding puts just past a knot (by ~ulp of x/h) is
// evaluated on the neighbouring piece, which agrees there to O(delta^2) in N
// (both kernels are C1) and O(delta) in N'.  Linear keeps bspline.
template <int INTERP, int J>
MPM_HD void bspline_node(float d, float& N, float& dN) {
#pragma clang fp contract(off)
    const float xa = fabsf(d);
    if (INTERP == MPM_INTERP_QUADRATIC_BSPLINE) {
        if (J == 1) {
            N = 0.75f - xa * xa;
            dN = -2.0f * d;
        } else {
            const float t = 1.5f - xa;
            N = 0.5f * (t * t);
            dN = d - (J == 0 ? 1.5f : -1.5f);
        }
    } else if (INTERP == MPM_INTERP_CUBIC_BSPLINE) {
        if (J == 1 || J == 2) {
            const float x2 = xa * xa;
            N = (0.5f * x2 * xa - x2) + (2.0f / 3.0f);
            dN = 1.5f * d * xa - 2.0f * d;
        } else {
            const float t = 2.0f - xa;
            N = (t * t * t) * (1.0f / 6.0f);  // within an ulp of the quotient
            dN = (2.0f * d - (J == 0 ? 2.0f : -2.0f)) - 0.5f * (d * xa);
        }
    } else {
        bspline<INTERP>(d, N, dN);
    }
}

template <int INTERP> struct Stencil;
template <> struct Stencil<MPM_INTERP_LINEAR> { static constexpr int W = 2; static constexpr float SHIFT = 0.0f; };
template <> struct Stencil<MPM_INTERP_QUADRATIC_BSPLINE> { static constexpr int W = 3; static constexpr float SHIFT = 0.5f; };
template <> struct Stencil<MPM_INTERP_CUBIC_BSPLINE> { static constexpr int W = 4; static constexpr float SHIFT = 1.0f; };

// Base cell of the stencil along one axis (also the sort key's cell):
// b = floor(u - shift).  Quadratic and cubic: u = fl(x * fl(1/h)), one
// product.  Linear: u = fl(x/h), the quotient the reference's range uses,
// and where that quotient rounded up onto an integer (x < h * floor(u),
// decided exactly by the sign of one fused multiply-add) b is one lower:
// the linear weight derivative jumps at the support boundary, so the two
// nodes must be the ones with 0 <= (x - x_b)/h < 1 exactly.
template <int INTERP>
MPM_HD float base_cell(float x, float h, float inv_h) {
    if (INTERP == MPM_INTERP_LINEAR) {
        float fb = floorf(x / h);
        if (fmaf(-h, fb, x) < 0.0f) fb -= 1.0f;
        return fb;
    }
    return floorf(x * inv_h - Stencil<INTERP>::SHIFT);
}

// Per-axis weights of one particle: base node b (fixed W-node stencil, b =
// base_cell), N and N'/h for nodes b..b+W-1 at d = (x - x_j) * fl(1/h) with
// x_j = origin + h j, the reference's node position (gd_pos,
// src/mpm_solver3d.cu:160).  The reference's range loop
// ceil(x/h - r)..floor(x/h + r) (src/mpm_solver3d.cu:251-255) selects the
// same nodes with non-zero weight or derivative, except that for quadratic
// and cubic a node at |d| = r - delta, delta <= 2 ulp(x/h), may be in one
// set only; N and N' vanish continuously there (N < delta^2, |N'| < delta).
template <int INTERP>
MPM_HD void axis_weights(float x, float origin, float h, float inv_h, int& base, float N[Stencil<INTERP>::W],
                         float dN[Stencil<INTERP>::W]) {
    constexpr int W = Stencil<INTERP>::W;
    float fb = base_cell<INTERP>(x, h, inv_h);
    fb = fminf(fmaxf(fb, -1.0e9f), 1.0e9f);
    base = (int)fb;
    float d[W];
#pragma unroll
    for (int j = 0; j < W; ++j) d[j] = (x - (origin + h * (fb + (float)j))) * inv_h;
    float n[4], dn[4];
    bspline_node<INTERP, 0>(d[0], n[0], dn[0]);
    bspline_node<INTERP, 1>(d[1], n[1], dn[1]);
    if constexpr (W > 2) bspline_node<INTERP, 2>(d[2], n[2], dn[2]);
    if constexpr (W > 3) bspline_node<INTERP, 3>(d[3], n[3], dn[3]);
#pragma unroll
    for (int j = 0; j < W; ++j) {
        N[j] = n[j];
        dN[j] = dn[j] * inv_h;
    }
}

// ------------------------------------------------------ particle cell key
// Same function as the oracle's cell_key: bit-exact by construction (the
// base cell of base_cell, then integer arithmetic).  width = stencil nodes
// per axis (2 = linear).
MPM_HD uint32_t cell_key(const SceneDev& S, float x, float y, float z, float shift, int width) {
    float p[3] = {x, y, z};
    int res[3] = {S.rx, S.ry, S.rz};
    int b[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float fb;  // base_cell
        if (width == 2) {
            fb = floorf(p[i] / S.h);
            if (fmaf(-S.h, fb, p[i]) < 0.0f) fb -= 1.0f;
        } else {
            fb = floorf(p[i] * S.inv_h - shift);
        }
        bool ok = (fb >= (float)(-(width - 1))) && (fb <= (float)(res[i] - 1));
        if (!ok) return S.escape_key;
        b[i] = (int)fb;
    }
    uint32_t blk = (uint32_t)((b[0] + 4) >> 2) +
                   (uint32_t)S.kbx * ((uint32_t)((b[1] + 4) >> 2) + (uint32_t)S.kby * (uint32_t)((b[2] + 4) >> 2));
    uint32_t cell = (uint32_t)((b[0] + 4) & 3) | ((uint32_t)((b[1] + 4) & 3) << 2) | ((uint32_t)((b[2] + 4) & 3) << 4);
    return blk * 64u + cell;
}

// ------------------------------------------------------ boundary collision
// applyBoundaryCollision, src/collision_object3d.cu:8-54 (walls of the grid
// box, Coulomb friction, unnormalised edge/corner normals).
MPM_HD void boundary_collision(const SceneDev& S, float px, float py, float pz, float& vx,
                                                   float& vy, float& vz) {
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
    bool hit = false;
    if (px <= S.ox + MPM_FLT_EPS) { hit = true; n0 = 1.0f; }
    if (px >= S.tx - MPM_FLT_EPS) { hit = true; n0 = -1.0f; }
    if (py <= S.oy + MPM_FLT_EPS) { hit = true; n1 = 1.0f; }
    if (py >= S.ty - MPM_FLT_EPS) { hit = true; n1 = -1.0f; }
    if (pz <= S.oz + MPM_FLT_EPS) { hit = true; n2 = 1.0f; }
    if (pz >= S.tz - MPM_FLT_EPS) { hit = true; n2 = -1.0f; }
    if (!hit) return;
    float vn = (vx * n0 + vy * n1) + vz * n2;
    if (vn >= MPM_FLT_EPS) return;
    float tx = vx - vn * n0, ty = vy - vn * n1, tz = vz - vn * n2;
    float nt = sqrtf((tx * tx + ty * ty) + tz * tz);
    if (nt <= -S.friction * vn + MPM_FLT_EPS) {
        vx = 0.0f; vy = 0.0f; vz = 0.0f;
    } else {
        float k = S.friction * vn;
        vx = tx + k * (tx / nt);
        vy = ty + k * (ty / nt);
        vz = tz + k * (tz / nt);
    }
}

// The same update from a folded grid, whose planes hold the mass and
// q = p + dt f with f the elastic force only (the fused APIC step: G2P reads
// v* alone).  v* = q / m - dt g e_y for m > FLT_EPSILON, else 0 (the
// reference's light nodes keep v* = 0), then the boundary collision.  In
// exact arithmetic this is grid_node_update's v*.
MPM_HD void grid_node_update_folded(const SceneDev& S, int i, int j, int k, float m, float qx, float qy, float qz,
                                    float& wx, float& wy, float& wz) {
    if (m > MPM_FLT_EPS) {
        const float im = 1.0f / m;
        wx = qx * im;
        wy = fmaf(-S.dt, S.gravity, qy * im);
        wz = qz * im;
    } else {
        wx = 0.0f; wy = 0.0f; wz = 0.0f;
    }
    float gx = (S.ox + S.h * (float)i) + S.dt * wx;
    float gy = (S.oy + S.h * (float)j) + S.dt * wy;
    float gz = (S.oz + S.h * (float)k) + S.dt * wz;
    boundary_collision(S, gx, gy, gz, wx, wy, wz);
}

// ------------------------------------------------------- grid addressing
// Storage node index (block-major 4x4x4, x fastest inside a block).
MPM_HD int64_t node_index(const SceneDev& S, int i, int j, int k) {
    int64_t sb = ((int64_t)(k >> 2) * S.nby + (j >> 2)) * S.nbx + (i >> 2);
    return sb * 64 + (((k & 3) << 4) | ((j & 3) << 2) | (i & 3));
}
MPM_HD int64_t storage_block(const SceneDev& S, int i, int j, int k) {
    return ((int64_t)(k >> 2) * S.nby + (j >> 2)) * S.nbx + (i >> 2);
}

// Grid velocity update of one node: computeGravityForces
// (src/mpm_solver3d.cu:394-403), CollocatedGridData3D::updateVelocity
// (src/grid3d.cu:36-41) and gridCollision (src/mpm_solver3d.cu:405-426).
// In: mass, momentum (p), elastic force (f).  Out: v (the reference's
// "velocity" field: p/m, or the raw momentum when m <= FLT_EPSILON) and v*.
MPM_HD void grid_node_update(const SceneDev& S, int i, int j, int k, float m, float px, float py,
                                                 float pz, float fx, float fy, float fz, float& vx, float& vy,
                                                 float& vz, float& wx, float& wy, float& wz) {
    bool heavy = m > MPM_FLT_EPS;
    if (heavy) {
        // one correctly rounded reciprocal instead of six divisions: each
        // quotient differs from p/m by at most one ulp
        fy = fmaf(-S.gravity, m, fy);
        const float im = 1.0f / m;
        vx = px * im; vy = py * im; vz = pz * im;
        wx = vx + (S.dt * fx) * im;
        wy = vy + (S.dt * fy) * im;
        wz = vz + (S.dt * fz) * im;
    } else {
        vx = px; vy = py; vz = pz;
        wx = 0.0f; wy = 0.0f; wz = 0.0f;
    }
    float gx = (S.ox + S.h * (float)i) + S.dt * wx;
    float gy = (S.oy + S.h * (float)j) + S.dt * wy;
    float gz = (S.oz + S.h * (float)k) + S.dt * wz;
    boundary_collision(S, gx, gy, gz, wx, wy, wz);
}
