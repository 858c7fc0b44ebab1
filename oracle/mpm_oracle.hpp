// mpm_oracle.hpp -- CPU ORACLE (test infrastructure only).
//
// Serial restatement of the reference's explicit MPM snow step, used as the
// checker for the HIP solver.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product never does.
//
// Real = float reproduces the reference's FP32 storage together with the
// FP64 promotions its C++ source implies (double literals, pow/exp); Real =
// double is the FP64 rerun of the same algorithm, used to measure the FP32
// drift.  Each function cites the reference file:line it restates (paths are
// relative to the reference repository root).  Loop orders follow the
// reference (x outer, z inner; particles in array order) so the serial sums
// are one valid ordering of the reference's atomicAdds.
//
// Pinning (what checks each part, see tests/):
//   weights         closed forms, partition of unity, finite differences
//   svd3            numpy.linalg.svd (LAPACK) + invariants
//   stress          closed forms (rotation, s*I) + finite difference of the
//                   reference's own energy (src/material_point3d.cu:73-92)
//   collision       hand-computed cases
//   initialize      lattice closed form V0 = h^3 / ppc
//   step            conservation, single-particle free fall, FP32 vs FP64
//   keys / sort     std::stable_sort (integer, exact)
//   density raster  closed forms (one particle on a node), lattice density
//                   ppc m / h^3, sum density h^3 = sum m, an independent
//                   numpy raster
// Threads: with Solver::threads > 1 (the all-cores CPU baseline only) the
// particle and node loops run under OpenMP and the grid sums become atomic
// adds, mirroring the reference's atomicAdd scatter; the order of those sums
// then varies between runs, so every parity check uses threads = 1, where
// each loop runs in array order as stated above.
// The third-party SVD the reference calls (3x3_SVD_CUDA, unpinned version,
// submodule absent) and Eigen's evaluation order are not reproducible
// bit-for-bit; parity with them is tolerance-level only.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "../include/mpm_hip.h"

namespace mpm_oracle {

static constexpr float kFltEps = std::numeric_limits<float>::epsilon();

// ------------------------------------------------------------------ types
template <class R>
struct V3 {
    R a[3];
    R& operator[](int i) { return a[i]; }
    const R& operator[](int i) const { return a[i]; }
};
template <class R>
struct M3 {
    R a[3][3];
    R* operator[](int i) { return a[i]; }
    const R* operator[](int i) const { return a[i]; }
};

template <class R> V3<R> v3(R x, R y, R z) { V3<R> v; v[0] = x; v[1] = y; v[2] = z; return v; }
template <class R> V3<R> vzero() { return v3<R>(0, 0, 0); }
template <class R> M3<R> mzero() { M3<R> m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = 0; return m; }
template <class R> M3<R> eye() { M3<R> m = mzero<R>(); m[0][0] = m[1][1] = m[2][2] = 1; return m; }

// Eigen fixed-size semantics: coefficient-wise ops in R; a 3-term dot is
// ((a0*b0 + a1*b1) + a2*b2).
template <class R> V3<R> add(const V3<R>& x, const V3<R>& y) { return v3<R>(x[0] + y[0], x[1] + y[1], x[2] + y[2]); }
template <class R> V3<R> sub(const V3<R>& x, const V3<R>& y) { return v3<R>(x[0] - y[0], x[1] - y[1], x[2] - y[2]); }
template <class R> V3<R> scale(R s, const V3<R>& x) { return v3<R>(s * x[0], s * x[1], s * x[2]); }
template <class R> R dot(const V3<R>& x, const V3<R>& y) { return (x[0] * y[0] + x[1] * y[1]) + x[2] * y[2]; }
template <class R> R norm(const V3<R>& x) { return std::sqrt(dot(x, x)); }
template <class R> M3<R> madd(const M3<R>& x, const M3<R>& y) { M3<R> m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = x[i][j] + y[i][j]; return m; }
template <class R> M3<R> msub(const M3<R>& x, const M3<R>& y) { M3<R> m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = x[i][j] - y[i][j]; return m; }
template <class R> M3<R> mscale(R s, const M3<R>& x) { M3<R> m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = s * x[i][j]; return m; }
template <class R> M3<R> tr(const M3<R>& x) { M3<R> m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = x[j][i]; return m; }
template <class R> M3<R> mm(const M3<R>& x, const M3<R>& y) {
    M3<R> m;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = (x[i][0] * y[0][j] + x[i][1] * y[1][j]) + x[i][2] * y[2][j];
    return m;
}
template <class R> V3<R> mv(const M3<R>& x, const V3<R>& y) {
    return v3<R>((x[0][0] * y[0] + x[0][1] * y[1]) + x[0][2] * y[2],
                 (x[1][0] * y[0] + x[1][1] * y[1]) + x[1][2] * y[2],
                 (x[2][0] * y[0] + x[2][1] * y[1]) + x[2][2] * y[2]);
}
template <class R> M3<R> outer(const V3<R>& x, const V3<R>& y) { M3<R> m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = x[i] * y[j]; return m; }
// Eigen 3x3 determinant: expansion along the first column.
template <class R> R det(const M3<R>& m) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[1][0] * (m[0][1] * m[2][2] - m[0][2] * m[2][1]) +
           m[2][0] * (m[0][1] * m[1][2] - m[0][2] * m[1][1]);
}
// Eigen 3x3 inverse: cofactors times 1/det (det from the first column).
template <class R> M3<R> inv(const M3<R>& m) {
    auto cof = [&](int i, int j) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
    };
    R c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    R d = (c00 * m[0][0] + c10 * m[1][0]) + c20 * m[2][0];
    R id = R(1) / d;
    M3<R> r;
    r[0][0] = c00 * id; r[0][1] = c10 * id; r[0][2] = c20 * id;
    r[1][0] = cof(0, 1) * id; r[1][1] = cof(1, 1) * id; r[1][2] = cof(2, 1) * id;
    r[2][0] = cof(0, 2) * id; r[2][1] = cof(1, 2) * id; r[2][2] = cof(2, 2) * id;
    return r;
}
template <class R> M3<R> diagm(const V3<R>& s) { M3<R> m = mzero<R>(); m[0][0] = s[0]; m[1][1] = s[1]; m[2][2] = s[2]; return m; }

// ------------------------------------------------------------------- SVD
// A = U diag(s) V^T with U, V rotations (det +1), s0 >= s1 >= |s2| and
// sign(s2) = sign(det A): the convention of the McAdams et al. 2011 method
// the reference calls (src/svd_eigen.cu:8-17).  Computed in FP64 with cyclic
// Jacobi on A^T A run to convergence, a descending sort of the columns of
// A V, and a Givens QR of A V; the FP32 oracle rounds the FP64 factors.
inline void svd3_f64(const double A[3][3], double U[3][3], double s[3], double V[3][3]) {
    double S[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        double dg = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
        if (off == 0.0 || off <= 1e-36 * dg) break;
        for (auto& pq : pairs) {
            int p = pq[0], q = pq[1];
            double apq = S[p][q];
            if (apq == 0.0) continue;
            double theta = (S[q][q] - S[p][p]) / (2.0 * apq);
            double t;
            if (std::fabs(theta) > 1e150) t = 0.5 / theta;
            else t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
            for (int k = 0; k < 3; ++k) {
                double skp = S[k][p], skq = S[k][q];
                S[k][p] = c * skp - sn * skq;
                S[k][q] = sn * skp + c * skq;
            }
            for (int k = 0; k < 3; ++k) {
                double spk = S[p][k], sqk = S[q][k];
                S[p][k] = c * spk - sn * sqk;
                S[q][k] = sn * spk + c * sqk;
            }
            S[p][q] = S[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {
                double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - sn * vkq;
                V[k][q] = sn * vkp + c * vkq;
            }
        }
    }
    double B[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i][j] = A[i][0] * V[0][j] + A[i][1] * V[1][j] + A[i][2] * V[2][j];
    auto colnorm = [&](int j) { return B[0][j] * B[0][j] + B[1][j] * B[1][j] + B[2][j] * B[2][j]; };
    auto swapneg = [&](int i, int j) {  // swap columns i,j and negate the new column j (det V stays +1)
        for (int k = 0; k < 3; ++k) {
            double t = B[k][i]; B[k][i] = B[k][j]; B[k][j] = -t;
            t = V[k][i]; V[k][i] = V[k][j]; V[k][j] = -t;
        }
    };
    if (colnorm(0) < colnorm(1)) swapneg(0, 1);
    if (colnorm(0) < colnorm(2)) swapneg(0, 2);
    if (colnorm(1) < colnorm(2)) swapneg(1, 2);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) U[i][j] = (i == j) ? 1.0 : 0.0;
    const int rot[3][3] = {{0, 1, 0}, {0, 2, 0}, {1, 2, 1}};  // rows p,q; column zeroed
    for (auto& g : rot) {
        int p = g[0], q = g[1], col = g[2];
        double a = B[p][col], b = B[q][col];
        double r = std::hypot(a, b);
        double c = 1.0, sn = 0.0;
        if (r > 0.0) { c = a / r; sn = b / r; }
        for (int k = 0; k < 3; ++k) {
            double bp = B[p][k], bq = B[q][k];
            B[p][k] = c * bp + sn * bq;
            B[q][k] = -sn * bp + c * bq;
        }
        B[q][col] = 0.0;
        for (int k = 0; k < 3; ++k) {
            double ukp = U[k][p], ukq = U[k][q];
            U[k][p] = c * ukp + sn * ukq;
            U[k][q] = -sn * ukp + c * ukq;
        }
    }
    s[0] = B[0][0]; s[1] = B[1][1]; s[2] = B[2][2];
}

template <class R>
void svd3(const M3<R>& A, M3<R>& U, V3<R>& s, M3<R>& V) {
    double a[3][3], u[3][3], sv[3], v[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = double(A[i][j]);
    svd3_f64(a, u, sv, v);
    for (int i = 0; i < 3; ++i) {
        s[i] = R(sv[i]);
        for (int j = 0; j < 3; ++j) { U[i][j] = R(u[i][j]); V[i][j] = R(v[i][j]); }
    }
}

// -------------------------------------------------- interpolation kernels
// src/interpolation3d.cu:88-142.  The reference mixes float arguments with
// double literals and pow(); those promotions are kept (for Real = float the
// products of two floats are still rounded to float, as in the reference).
template <class R> R linearN(R x) {             // :88-93
    R xabs = std::fabs(x);
    if (xabs < 1.0) return R(1.0 - xabs);
    return R(0.0);
}
template <class R> R linearNDot(R x) {          // :96-102
    R xabs = std::fabs(x);
    if (xabs < 1.0 && xabs > R(kFltEps)) return -x / xabs;
    return R(0.0);
}
template <class R> R quadN(R x) {               // :105-112
    R xabs = std::fabs(x);
    if (xabs < 0.5) return R(0.75 - xabs * xabs);
    if (xabs < 1.5) return R(0.5 * std::pow(1.5 - xabs, 2));
    return R(0.0);
}
template <class R> R quadNDot(R x) {            // :115-122
    R xabs = std::fabs(x);
    if (xabs < 0.5) return R(-2.0 * x);
    if (xabs < 1.5) return R(-(1.5 * x / xabs) + x);
    return R(0.0);
}
template <class R> R cubicN(R x) {              // :125-132
    R xabs = std::fabs(x);
    if (xabs < 1.0) return R(0.5 * std::pow(double(xabs), 3) - xabs * xabs + 2.0 / 3.0);
    if (xabs < 2.0) return R(std::pow(2.0 - xabs, 3) / 6.0);
    return R(0.0);
}
template <class R> R cubicNDot(R x) {           // :135-142
    R xabs = std::fabs(x);
    if (xabs < 1.0) return R(1.5 * x * xabs - 2.0 * x);
    if (xabs < 2.0) return R(-(2.0 * x / xabs) + 2.0 * x - 0.5 * (x * xabs));
    return R(0.0);
}

template <class R>
struct Interpolator {                            // include/interpolation3d.h:11-56
    int type;
    R range;
    explicit Interpolator(int t) : type(t) {     // src/interpolation3d.cu:69-85
        range = (t == MPM_INTERP_LINEAR) ? R(1.0) : (t == MPM_INTERP_QUADRATIC_BSPLINE) ? R(1.5) : R(2.0);
    }
    R N(R x) const {
        if (type == MPM_INTERP_LINEAR) return linearN(x);
        if (type == MPM_INTERP_QUADRATIC_BSPLINE) return quadN(x);
        return cubicN(x);
    }
    R Nd(R x) const {
        if (type == MPM_INTERP_LINEAR) return linearNDot(x);
        if (type == MPM_INTERP_QUADRATIC_BSPLINE) return quadNDot(x);
        return cubicNDot(x);
    }
    R weight3D(const V3<R>& p, const V3<R>& g, R h) const {            // :14-33
        R x = (p[0] - g[0]) / h, y = (p[1] - g[1]) / h, z = (p[2] - g[2]) / h;
        return N(x) * N(y) * N(z);
    }
    V3<R> weightGradient3D(const V3<R>& p, const V3<R>& g, R h) const { // :36-66
        R x = (p[0] - g[0]) / h, y = (p[1] - g[1]) / h, z = (p[2] - g[2]) / h;
        return v3<R>((Nd(x) * N(y) * N(z)) / h, (N(x) * Nd(y) * N(z)) / h, (N(x) * N(y) * Nd(z)) / h);
    }
};

// ---------------------------------------------------------------- particle
template <class R>
struct Particle {  // include/material_point3d.h:15-21 + include/apic_material_point3d.h:23
    R volume0, mass;
    V3<R> position, velocity;
    M3<R> Fe, Fp;
    R mu0, lambda0, hardening, crit_c, crit_s;
    int model;
    M3<R> B, D, Dinv;
};

// MaterialPoint3D constructor, src/material_point3d.cu:10-30 (APIC adds
// B = D = Dinv = 0, src/apic_material_point3d.cu:31-33).  E*nu is a float
// product in the reference; the rest is evaluated in double.
template <class R>
Particle<R> make_particle(const float* x, const float* v, float m, float xi, float E, float nu, float thc,
                          float ths, int model) {
    Particle<R> p;
    p.position = v3<R>(x[0], x[1], x[2]);
    p.velocity = v3<R>(v[0], v[1], v[2]);
    p.mass = m;
    p.hardening = xi;
    p.crit_c = thc;
    p.crit_s = ths;
    p.model = model;
    R Er = E, nur = nu;
    p.mu0 = R(Er / (2.0 * (1.0 + nur)));
    p.lambda0 = R((Er * nur) / ((1.0 + nur) * (1.0 - 2.0 * nur)));
    p.Fe = eye<R>();
    p.Fp = eye<R>();
    p.B = mzero<R>();
    p.D = mzero<R>();
    p.Dinv = mzero<R>();
    p.volume0 = 0;
    return p;
}

// plasticLame, src/material_point3d.cu:158-162 (exp in double, factor
// stored as float).
template <class R>
void plastic_lame(const Particle<R>& p, R& mu, R& la) {
    R Jp = det(p.Fp);
    R factor = R(std::exp(p.hardening * (1.0 - Jp)));
    mu = p.mu0 * factor;
    la = p.lambda0 * factor;
}

// fixedCorotatedEnergyDerivativeTimesDeformationGradientTranspose,
// src/material_point3d.cu:110-128: 2 mu (Fe - R) Fe^T + lambda (Je-1) Je I.
template <class R>
M3<R> fixed_corotated_PFt(const Particle<R>& p) {
    R mu, la;
    plastic_lame(p, mu, la);
    M3<R> U, V;
    V3<R> s;
    svd3(p.Fe, U, s, V);
    M3<R> Rm = mm(U, tr(V));
    R Je = det(p.Fe);
    M3<R> a = mm(mscale(R(2.0 * mu), msub(p.Fe, Rm)), tr(p.Fe));
    M3<R> b = mscale(R(la * (Je - 1.0) * Je), eye<R>());
    return madd(a, b);
}

// NeoHookeanEnergyEnergyDerivativeTimesDeformationGradientTranspose,
// src/material_point3d.cu:131-155.  Note the side effect: Fe is replaced by
// its projection with singular values >= FLT_EPSILON (:142).
template <class R>
M3<R> neo_hookean_PFt(Particle<R>& p) {
    R mu, la;
    plastic_lame(p, mu, la);
    M3<R> U, V;
    V3<R> s;
    svd3(p.Fe, U, s, V);
    for (int i = 0; i < 3; ++i) s[i] = std::max(s[i], R(kFltEps));
    p.Fe = mm(mm(U, diagm(s)), tr(V));
    R Je = det(p.Fe);
    M3<R> a = mscale(mu, msub(mm(p.Fe, tr(p.Fe)), eye<R>()));
    M3<R> b = mscale(la * R(std::log(Je)), eye<R>());
    return madd(a, b);
}

// volumeTimesCauchyStress, src/material_point3d.cu:33-45.
template <class R>
M3<R> volume_times_cauchy_stress(Particle<R>& p) {
    if (p.model == MPM_MODEL_NEO_HOOKEAN) return mscale(p.volume0, neo_hookean_PFt(p));
    return mscale(p.volume0, fixed_corotated_PFt(p));
}

// updateDeformationGradient, src/material_point3d.cu:53-70.
template <class R>
void update_deformation_gradient(Particle<R>& p, const M3<R>& grad_v, R dt) {
    p.Fe = mm(madd(eye<R>(), mscale(dt, grad_v)), p.Fe);
    M3<R> F = mm(p.Fe, p.Fp);
    M3<R> U, V;
    V3<R> s;
    svd3(p.Fe, U, s, V);
    R hi = R(1.0f) + p.crit_s, lo = R(1.0f) - p.crit_c;
    for (int i = 0; i < 3; ++i) s[i] = std::max(std::min(s[i], hi), lo);
    M3<R> S = diagm(s);
    p.Fe = mm(mm(U, S), tr(V));
    p.Fp = mm(mm(mm(V, inv(S)), tr(U)), F);
}

// applyBoundaryCollision, src/collision_object3d.cu:8-54.
template <class R>
V3<R> apply_boundary_collision(const V3<R>& pos, const V3<R>& vel, const V3<R>& g000, const V3<R>& g111, R mu_b) {
    V3<R> updated = vel, rel = vel, normal = vzero<R>();
    bool collided = false;
    for (int i = 0; i < 3; ++i) {
        if (pos[i] <= g000[i] + R(kFltEps)) { collided = true; normal[i] = 1; }
        if (pos[i] >= g111[i] - R(kFltEps)) { collided = true; normal[i] = -1; }
    }
    if (collided) {
        R vn = dot(rel, normal);
        if (vn >= R(kFltEps)) return updated;
        V3<R> vt = sub(rel, scale(vn, normal));
        R nt = norm(vt);
        if (nt <= -mu_b * vn + R(kFltEps)) {
            rel = vzero<R>();
        } else {
            V3<R> vhat = v3<R>(vt[0] / nt, vt[1] / nt, vt[2] / nt);
            rel = add(vt, scale(mu_b * vn, vhat));
        }
        updated = rel;
    }
    return updated;
}

// Grid sum: an atomic add when particle loops run on several threads.
template <class R>
inline void grid_add(R& dst, R v, bool atomic) {
    if (atomic) {
#pragma omp atomic
        dst += v;
    } else {
        dst += v;
    }
}

// -------------------------------------------------------------------- grid
template <class R>
struct Node {  // CollocatedGridData3D, include/grid3d.h:27-47
    int index[3];
    V3<R> force, velocity, velocity_star;
    R mass;
};

// ------------------------------------------------------------------ solver
template <class R>
struct Solver {
    mpm_desc desc;
    bool apic;
    Interpolator<R> interp;
    V3<R> origin, target;
    R h, mu_b, blend, dt, gravity;
    int res[3];
    bool particle_collision;
    std::vector<Particle<R>> P;
    std::vector<Node<R>> G;
    int64_t steps = 0;
    int threads = 1;  // > 1: OpenMP loops, atomic grid sums (timing only)

    Solver(const mpm_desc& d, std::vector<Particle<R>> parts)
        : desc(d), apic(d.transfer == MPM_TRANSFER_APIC), interp(d.interp), P(std::move(parts)) {
        for (int i = 0; i < 3; ++i) res[i] = d.resolution[i];
        h = d.stride;
        mu_b = d.friction;
        blend = d.blend;
        dt = d.dt;
        gravity = d.gravity;
        particle_collision = d.particle_collision != 0;
        origin = v3<R>(d.origin[0], d.origin[1], d.origin[2]);
        // Grid3DSettings: target = origin + stride * resolution, src/grid3d.cu:8-16
        target = add(origin, v3<R>(h * R(res[0]), h * R(res[1]), h * R(res[2])));
        size_t nn = size_t(res[0]) * res[1] * res[2];
        G.resize(nn);
        // node index tabulation, src/mpm_solver3d.cu:19-34
        for (size_t i = 0; i < nn; ++i) {
            G[i].index[0] = int(i % res[0]);
            G[i].index[1] = int((i / res[0]) % res[1]);
            G[i].index[2] = int((i / res[0]) / res[1]);
        }
        reset_grid();
    }

    int64_t node_id(int x, int y, int z) const { return x + int64_t(res[0]) * (y + int64_t(res[1]) * z); }
    V3<R> node_pos(int x, int y, int z) const {
        return add(origin, v3<R>(h * R(x), h * R(y), h * R(z)));
    }
    // Stencil bounds, src/mpm_solver3d.cu:251-255 (origin ignored, as in
    // the reference).
    void range(const V3<R>& x, int lo[3], int hi[3]) const {
        for (int i = 0; i < 3; ++i) {
            lo[i] = int(std::ceil(x[i] / h - interp.range));
            hi[i] = int(std::floor(x[i] / h + interp.range));
        }
    }
    bool inside(int x, int y, int z) const {
        return !(x < 0 || x >= res[0] || y < 0 || y >= res[1] || z < 0 || z >= res[2]);
    }

    // initialize(): P2GMass + G2PVolume, src/mpm_solver3d.cu:131-223;
    // APIC D/Dinv, src/apic_mpm_solver3d.cu:151-157.
    void initialize() {
        const int64_t np = int64_t(P.size());
        const bool mt = threads > 1;
#pragma omp parallel for schedule(static) if (mt) num_threads(threads)
        for (int64_t ip = 0; ip < np; ++ip) {
            Particle<R>& p = P[ip];
            int lo[3], hi[3];
            range(p.position, lo, hi);
            if (apic) {
                R k = 0;
                if (desc.interp == MPM_INTERP_CUBIC_BSPLINE) k = R(1 / 3.0);
                if (desc.interp == MPM_INTERP_QUADRATIC_BSPLINE) k = R(1 / 4.0);
                p.D = mscale(k * h * h, eye<R>());
                p.Dinv = inv(p.D);
            }
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) {
                        if (!inside(x, y, z)) continue;
                        R w = interp.weight3D(p.position, node_pos(x, y, z), h);
                        grid_add(G[node_id(x, y, z)].mass, p.mass * w, mt);
                    }
        }
#pragma omp parallel for schedule(static) if (mt) num_threads(threads)
        for (int64_t ip = 0; ip < np; ++ip) {
            Particle<R>& p = P[ip];
            int lo[3], hi[3];
            range(p.position, lo, hi);
            R density = 0;
            R inv_cell_vol = R(1.0 / (h * h * h));
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) {
                        if (!inside(x, y, z)) continue;
                        R w = interp.weight3D(p.position, node_pos(x, y, z), h);
                        R m_ip = G[node_id(x, y, z)].mass * w;
                        density += m_ip * inv_cell_vol;
                    }
            p.volume0 = p.mass / density;
        }
    }

    void reset_grid() {  // src/mpm_solver3d.cu:225-235, src/grid3d.cu:28-33
        const int64_t nn = int64_t(G.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t i = 0; i < nn; ++i) {
            Node<R>& n = G[i];
            n.mass = 0;
            n.force = vzero<R>();
            n.velocity = vzero<R>();
            n.velocity_star = vzero<R>();
        }
    }

    // particlesToGrid: PIC-FLIP src/mpm_solver3d.cu:237-294,
    // APIC src/apic_mpm_solver3d.cu:245-304.
    void particles_to_grid() {
        const int64_t np = int64_t(P.size());
        const bool mt = threads > 1;
#pragma omp parallel for schedule(static) if (mt) num_threads(threads)
        for (int64_t ip = 0; ip < np; ++ip) {
            Particle<R>& p = P[ip];  // the Neo-Hookean stress writes Fe back (src/material_point3d.cu:142)
            int lo[3], hi[3];
            range(p.position, lo, hi);
            M3<R> C = apic ? mm(p.B, p.Dinv) : mzero<R>();
            M3<R> vol_stress = mscale(R(-1), volume_times_cauchy_stress(p));
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) {
                        if (!inside(x, y, z)) continue;
                        V3<R> gp = node_pos(x, y, z);
                        R w = interp.weight3D(p.position, gp, h);
                        V3<R> gw = interp.weightGradient3D(p.position, gp, h);
                        R m_ip = p.mass * w;
                        V3<R> f_ip = mv(vol_stress, gw);
                        Node<R>& n = G[node_id(x, y, z)];
                        grid_add(n.mass, m_ip, mt);
                        if (apic) {
                            V3<R> v_ip = scale(m_ip, add(p.velocity, mv(C, sub(gp, p.position))));
                            for (int i = 0; i < 3; ++i) grid_add(n.velocity[i], v_ip[i], mt);
                        } else {
                            for (int i = 0; i < 3; ++i) grid_add(n.velocity[i], p.velocity[i] * m_ip, mt);
                        }
                        for (int i = 0; i < 3; ++i) grid_add(n.force[i], f_ip[i], mt);
                    }
        }
    }

    // computeGravityForces, src/mpm_solver3d.cu:394-403 (double product).
    void compute_gravity_forces() {
        const int64_t nn = int64_t(G.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t i = 0; i < nn; ++i) {
            Node<R>& n = G[i];
            if (n.mass > R(kFltEps)) n.force[1] = R(double(n.force[1]) - double(gravity) * double(n.mass));
        }
    }

    // updateGridVelocities -> CollocatedGridData3D::updateVelocity,
    // src/mpm_solver3d.cu:300-310, src/grid3d.cu:36-41.
    void update_grid_velocities(R dt_) {
        const int64_t nn = int64_t(G.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t i = 0; i < nn; ++i) {
            Node<R>& n = G[i];
            if (n.mass > R(kFltEps)) {
                for (int i = 0; i < 3; ++i) n.velocity[i] = n.velocity[i] / n.mass;
                for (int i = 0; i < 3; ++i) n.velocity_star[i] = n.velocity[i] + (dt_ * n.force[i]) / n.mass;
            }
        }
    }

    // gridCollision, src/mpm_solver3d.cu:405-426.
    void grid_collision(R dt_) {
        const int64_t nn = int64_t(G.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t i = 0; i < nn; ++i) {
            Node<R>& n = G[i];
            V3<R> pos = add(node_pos(n.index[0], n.index[1], n.index[2]), scale(dt_, n.velocity_star));
            n.velocity_star = apply_boundary_collision(pos, n.velocity_star, origin, target, mu_b);
        }
    }

    // gridToParticles: PIC-FLIP src/mpm_solver3d.cu:316-379,
    // APIC src/apic_mpm_solver3d.cu:326-388.
    void grid_to_particles(R dt_, R blend_) {
        const int64_t np = int64_t(P.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t ip = 0; ip < np; ++ip) {
            Particle<R>& p = P[ip];
            int lo[3], hi[3];
            range(p.position, lo, hi);
            M3<R> grad = mzero<R>();
            V3<R> vpic = vzero<R>(), vflip = p.velocity;
            if (apic) p.B = mzero<R>();
            for (int x = lo[0]; x <= hi[0]; ++x)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int z = lo[2]; z <= hi[2]; ++z) {
                        if (!inside(x, y, z)) continue;
                        V3<R> gp = node_pos(x, y, z);
                        R w = interp.weight3D(p.position, gp, h);
                        V3<R> gw = interp.weightGradient3D(p.position, gp, h);
                        const Node<R>& n = G[node_id(x, y, z)];
                        if (apic) {
                            V3<R> v_ip = scale(w, n.velocity_star);
                            vpic = add(vpic, v_ip);
                            grad = madd(grad, outer(n.velocity_star, gw));
                            p.B = madd(p.B, outer(v_ip, sub(gp, p.position)));
                        } else {
                            vpic = add(vpic, scale(w, n.velocity_star));
                            vflip = add(vflip, scale(w, sub(n.velocity_star, n.velocity)));
                            grad = madd(grad, outer(n.velocity_star, gw));
                        }
                    }
            V3<R> vel;
            if (apic) vel = vpic;
            else vel = add(scale(blend_, vflip), scale(R(1.0 - blend_), vpic));
            update_deformation_gradient(p, grad, dt_);
            p.velocity = vel;
        }
    }

    void particles_collision(R dt_) {  // src/mpm_solver3d.cu:428-445
        const int64_t np = int64_t(P.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t ip = 0; ip < np; ++ip) {
            Particle<R>& p = P[ip];
            p.velocity = apply_boundary_collision(add(p.position, scale(dt_, p.velocity)), p.velocity, origin, target, mu_b);
        }
    }

    void advect_particles(R dt_) {  // src/mpm_solver3d.cu:381-391, src/material_point3d.cu:48-50
        const int64_t np = int64_t(P.size());
#pragma omp parallel for schedule(static) if (threads > 1) num_threads(threads)
        for (int64_t ip = 0; ip < np; ++ip) P[ip].position = add(P[ip].position, scale(dt_, P[ip].velocity));
    }

    // simulateOneStep, src/mpm_solver3d.cu:100-125 / src/apic_mpm_solver3d.cu:99-124.
    void step() {
        reset_grid();
        particles_to_grid();
        compute_gravity_forces();
        update_grid_velocities(dt);
        grid_collision(dt);
        grid_to_particles(dt, blend);
        if (particle_collision) particles_collision(dt);
        advect_particles(dt);
        ++steps;
    }
};

// --------------------------------------------------------- density raster
// writeToOpenVDB's raster, src/mpm_solver3d.cu:510-563 (host code): for each
// particle in array order and each node of its stencil range (origin
// ignored, as in P2G) inside [0, res), density += mass * w * cell_vol_inv
// with cell_vol_inv = 1.0 / (h*h*h) rounded to R, w from the grid node's
// world position.  out: res_x*res_y*res_z values in node order
// x + rx (y + ry z), 0 where no particle reaches (the grid's background);
// active (optional) flags the nodes the reference's accessor writes.
template <class R>
void density_raster(const mpm_desc& d, const float* pos, const float* mass, int64_t n, std::vector<R>& out,
                    std::vector<uint8_t>* active) {
    const Interpolator<R> interp(d.interp);
    const int res[3] = {d.resolution[0], d.resolution[1], d.resolution[2]};
    const R h = R(d.stride);
    const R rng = interp.range;
    const V3<R> ori = v3<R>(R(d.origin[0]), R(d.origin[1]), R(d.origin[2]));
    const R cell_vol_inv = R(1.0 / (h * h * h));
    const size_t nn = size_t(res[0]) * res[1] * res[2];
    out.assign(nn, R(0));
    if (active) active->assign(nn, 0);
    for (int64_t p = 0; p < n; ++p) {
        const V3<R> x = v3<R>(R(pos[3 * p]), R(pos[3 * p + 1]), R(pos[3 * p + 2]));
        int lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
            lo[i] = int(std::ceil(x[i] / h - rng));
            hi[i] = int(std::floor(x[i] / h + rng));
        }
        for (int gx = lo[0]; gx <= hi[0]; ++gx)
            for (int gy = lo[1]; gy <= hi[1]; ++gy)
                for (int gz = lo[2]; gz <= hi[2]; ++gz) {
                    if (gx < 0 || gx >= res[0] || gy < 0 || gy >= res[1] || gz < 0 || gz >= res[2]) continue;
                    const size_t id = size_t(gx) + size_t(res[0]) * (size_t(gy) + size_t(res[1]) * size_t(gz));
                    const V3<R> gp = add(ori, v3<R>(h * R(gx), h * R(gy), h * R(gz)));
                    const R w = interp.weight3D(x, gp, h);
                    out[id] = out[id] + R(mass[p]) * w * cell_vol_inv;
                    if (active) (*active)[id] = 1;
                }
    }
}

// ------------------------------------------------ particle-by-cell sort key
// No reference counterpart (the reference never sorts).  The key of a
// particle is its base cell b = floor(fl(x/h) - s) per axis (s = 0, 0.5, 1
// for linear, quadratic, cubic), grouped into 4x4x4 cell blocks shifted by
// one block so b >= -3 maps to a block index >= 0:
//   key = block_linear(((b+4)>>2)) * 64 + in-block cell ((b+4)&3)
// A particle whose stencil misses the grid on some axis (b outside
// [-(S-1), res-1], or x not finite) gets the escape key, one past the largest
// block key.  Evaluated in float as on the device: b = floor(x * fl(1/h) - s)
// (one rounded product); linear: b = floor(fl(x/h)), lowered by one when
// x < h * b exactly (the quotient rounded up onto an integer).
struct KeyParams {
    int res[3];
    int kbn[3];
    float h;
    float inv_h;
    float shift;
    int width;
    uint32_t escape;
};
inline KeyParams key_params(const mpm_desc& d) {
    KeyParams k;
    for (int i = 0; i < 3; ++i) {
        k.res[i] = d.resolution[i];
        k.kbn[i] = ((d.resolution[i] + 3) >> 2) + 1;
    }
    k.h = d.stride;
    volatile float one = 1.0f;
    k.inv_h = one / d.stride;
    k.shift = d.interp == MPM_INTERP_LINEAR ? 0.0f : d.interp == MPM_INTERP_QUADRATIC_BSPLINE ? 0.5f : 1.0f;
    k.width = d.interp == MPM_INTERP_LINEAR ? 2 : d.interp == MPM_INTERP_QUADRATIC_BSPLINE ? 3 : 4;
    k.escape = uint32_t(k.kbn[0]) * uint32_t(k.kbn[1]) * uint32_t(k.kbn[2]) * 64u;
    return k;
}
inline uint32_t cell_key(const KeyParams& k, const float* x) {
    int b[3];
    for (int i = 0; i < 3; ++i) {
        float fb;
        if (k.width == 2) {
            volatile float u = x[i] / k.h;
            fb = std::floor(u);
            if (std::fma(-k.h, fb, x[i]) < 0.0f) fb -= 1.0f;
        } else {
            volatile float u = x[i] * k.inv_h;
            volatile float t = u - k.shift;
            fb = std::floor(t);
        }
        if (!(fb >= float(-(k.width - 1)) && fb <= float(k.res[i] - 1))) return k.escape;
        b[i] = int(fb);
    }
    uint32_t blk = uint32_t((b[0] + 4) >> 2) +
                   uint32_t(k.kbn[0]) * (uint32_t((b[1] + 4) >> 2) + uint32_t(k.kbn[1]) * uint32_t((b[2] + 4) >> 2));
    uint32_t cell = uint32_t((b[0] + 4) & 3) | (uint32_t((b[1] + 4) & 3) << 2) | (uint32_t((b[2] + 4) & 3) << 4);
    return blk * 64u + cell;
}

}  // namespace mpm_oracle
