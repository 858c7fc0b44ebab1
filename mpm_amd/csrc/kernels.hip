// kernels.hip -- particle/grid kernels of the MI355X explicit MPM step.
//
// Work decomposition: particles are sorted by cell (key = 4x4x4 block,
// then cell in block); a work item is a run of <= kItemMax sorted particles
// of one block.  P2G and G2P are persistent kernels: each workgroup walks
// the item list and stages the block's node tile (the block plus the
// stencil halo, T = 4 + W - 1 nodes per axis) in LDS:
//   P2G: particles scatter mass, momentum and force into the LDS tile (cell
//        sums over a wave-level segmented scan, one plain LDS update per
//        cell); the tile is flushed to HBM with one global float atomic per
//        non-zero node value.
//   G2P: the tile is loaded once per item; in the fused step the grid
//        velocity update (gravity, p/m, v* = v + dt f/m, boundary collision)
//        is computed while loading, so there is no separate grid pass.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "blockscan.h"
#include "devmath.h"
#include "launch.h"

namespace mpm {

constexpr int kG2PWG = 64;  // default G2P workgroup: one wave per item

// Persistent launches are sized to one resident round: the occupancy of
// the kernel instance (VGPRs, LDS, workgroup size) times the compute units.
// A larger grid leaves a partial second round (a tail), a smaller one idles
// SIMD slots.  requested > 0 overrides.  Cached per (kernel, device).
static int device_cus() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    static int cus[64] = {0};
    if (dev < 0 || dev >= 64) return 1;
    if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus[dev] = 1;
    return cus[dev] > 0 ? cus[dev] : 1;
}
template <typename F>
static unsigned resident_grid(F* kernel, int block, int requested) {
    if (requested > 0) return (unsigned)requested;
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_pair((const void*)kernel, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return (unsigned)it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1) {
        (void)hipGetLastError();
        nb = 1;
    }
    const int grid = nb * device_cus();
    cache[key] = grid;
    return (unsigned)grid;
}
#define MPM_PERSISTENT(K, BLOCK, REQ, ST, ...) \
    hipLaunchKernelGGL(K, dim3(resident_grid(K, BLOCK, REQ)), dim3(BLOCK), 0, ST, __VA_ARGS__)

// Minimum waves per SIMD asked of the register allocator for G2P.  The
// linear and quadratic G2P need 147-161 VGPRs; capped at 128 (4 waves per
// SIMD) they spill 60-90 bytes per lane, which costs more than the fourth
// wave gains.  Measured with 128-thread workgroups, per step: the 256-scene
// APIC batch 2.54 -> 2.39 ms, the same scenes as PIC-FLIP 2.86 -> 2.51 ms,
// the 8.4M-particle slab's two-kernel step 1.52 -> 1.44 ms, going from 4
// waves to 3 (168 VGPRs, no spill).  PIC-FLIP G2P keeps the compiler's
// allocation for the cubic stencil and 256 threads (MPM_PIC_G2P_WAVES; cubic
// 156-179 VGPRs).
#ifndef MPM_APIC_G2P_WAVES
#define MPM_APIC_G2P_WAVES 3
#endif
#ifndef MPM_APIC_G2P_WAVES_WG128
#define MPM_APIC_G2P_WAVES_WG128 3
#endif
#ifndef MPM_PIC_G2P_WAVES
#define MPM_PIC_G2P_WAVES 1
#endif
#ifndef MPM_PIC_G2P_WAVES_LQ
#define MPM_PIC_G2P_WAVES_LQ 3
#endif
template <int INTERP, int APIC, int WG>
struct G2PWaves {
    static constexpr int value =
        APIC ? ((WG == 128 && INTERP != MPM_INTERP_CUBIC_BSPLINE) ? MPM_APIC_G2P_WAVES_WG128 : MPM_APIC_G2P_WAVES)
             : ((WG <= 128 && INTERP != MPM_INTERP_CUBIC_BSPLINE) ? MPM_PIC_G2P_WAVES_LQ : MPM_PIC_G2P_WAVES);
};

// Minimum waves per SIMD for P2G.  The folded (4-value) and mass-only
// P2G of the linear and quadratic stencils fit 96 VGPRs without spills
// (the folded P2G since its node values are regrouped as yz R(a) + Nx[a]
// Syz(b,c)), so they run 5 waves per SIMD (measured: the two-kernel slab
// step 1.412 -> 1.398 ms and the 256-scene APIC batch 2.354 -> 2.313 ms
// from 4 waves; 3 -> 4 waves had taken APIC P2G from 0.674 to 0.616 ms).
// The 7-value P2G and the cubic stencil keep the compiler's allocation.
#ifndef MPM_P2G_WAVES
#define MPM_P2G_WAVES 5
#endif
template <int INTERP, int MODE>
struct P2GWaves {
    static constexpr int value = (MODE == 0 || INTERP == MPM_INTERP_CUBIC_BSPLINE) ? 1 : MPM_P2G_WAVES;
};

// ------------------------------------------------------------------ keys
// Cell key of each particle, plus the scene's key offset in a batch.
template <int INTERP>
__global__ void k_keys(SceneDev S, ParticlesDev P, uint32_t key_offset, uint32_t* __restrict__ keys) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int64_t N = P.n;
    keys[i] = key_offset +
              cell_key(S, P.x[i], P.x[N + i], P.x[2 * N + i], Stencil<INTERP>::SHIFT, Stencil<INTERP>::W);
}

void launch_keys(const SceneDev& S, const ParticlesDev& P, int interp, uint32_t* keys, hipStream_t st,
                 uint32_t key_offset) {
    if (P.n <= 0) return;
    dim3 g((unsigned)((P.n + 255) / 256)), b(256);
    switch (interp) {
        case MPM_INTERP_LINEAR:
            hipLaunchKernelGGL(k_keys<MPM_INTERP_LINEAR>, g, b, 0, st, S, P, key_offset, keys);
            break;
        case MPM_INTERP_QUADRATIC_BSPLINE:
            hipLaunchKernelGGL(k_keys<MPM_INTERP_QUADRATIC_BSPLINE>, g, b, 0, st, S, P, key_offset, keys);
            break;
        default: hipLaunchKernelGGL(k_keys<MPM_INTERP_CUBIC_BSPLINE>, g, b, 0, st, S, P, key_offset, keys); break;
    }
}

__global__ void k_iota(uint32_t* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}
void launch_iota(uint32_t* p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

// ----------------------------------------------------------------- items
// A work item starts where the key block changes (in a batch this includes
// every scene boundary, since scene key ranges are 64-aligned) and at
// every multiple of item_max within a block.
// The head flags of a thread's IPT consecutive sorted keys from base: the
// keys in 16-byte loads and the key before base.
template <int IPT>
__device__ __forceinline__ void item_heads(const uint32_t* __restrict__ sk, int64_t n, int64_t base,
                                           uint32_t item_max, uint32_t (&key)[IPT], bool (&head)[IPT]) {
    static_assert(IPT % 4 == 0, "keys are read four at a time");
    if (base + IPT <= n) {
#pragma unroll
        for (int j = 0; j < IPT; j += 4) {
            const uint4 k4 = *reinterpret_cast<const uint4*>(sk + base + j);  // base is a multiple of IPT
            key[j] = k4.x; key[j + 1] = k4.y; key[j + 2] = k4.z; key[j + 3] = k4.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < IPT; ++j) key[j] = base + j < n ? sk[base + j] : 0u;
    }
    uint32_t prev = base > 0 && base <= n ? sk[base - 1] : 0u;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const int64_t i = base + j;
        head[j] = i < n && (i == 0 || (key[j] >> 6) != (prev >> 6) || ((uint64_t)i % item_max) == 0);
        prev = key[j];
    }
}
// Work items in two passes over the sorted keys, with the head flags
// recomputed in each pass instead of stored: per-tile head counts, a scan
// of the tile counts, then every head written at its rank.  items[k] is the
// start of item k, items[n + k] its scene-local key block and, for a batch,
// item_scene[k] its scene (the last scene whose key offset is <= the item's
// first key).  plan (the incremental sort's words, or null): the sort of the
// step is done, so the first workgroup counts an incremental sort and
// resets the move counter for the step's G2P (and marks the snapshot that
// the step's transfer takes as valid).
template <int IPT>
__global__ void __launch_bounds__(kScanWG) k_item_count(const uint32_t* __restrict__ sk, int64_t n,
                                                        uint32_t item_max, uint32_t* __restrict__ partial,
                                                        uint32_t* __restrict__ plan) {
    __shared__ uint32_t ws[kScanWG / 64];
    if (plan && blockIdx.x == 0) {
        if (threadIdx.x < kMoveLists) plan[kPlanLists + threadIdx.x * kMoveCountStride] = 0u;
        if (threadIdx.x == 0) {
            plan[kPlanIncSteps] += plan_incremental(plan) ? 1u : 0u;
            plan[kPlanSnapOk] = 1u;
            plan[kPlanOverflow] = 0u;
        }
    }
    const int64_t base = ((int64_t)blockIdx.x * kScanWG + threadIdx.x) * IPT;
    uint32_t key[IPT];
    bool head[IPT];
    item_heads<IPT>(sk, n, base, item_max, key, head);
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) c += head[j] ? 1u : 0u;
    uint32_t tot;
    block_exclusive_scan(c, ws, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}
// tile_start: the scanned tile counts, or with RAW the raw counts, which
// each workgroup sums over the tiles before it (and the last one writes the
// item count): few tiles, one launch fewer.
template <bool RAW, int IPT>
__global__ void __launch_bounds__(kScanWG) k_item_emit(const uint32_t* __restrict__ sk, int64_t n, uint32_t item_max,
                                                       const uint32_t* __restrict__ tile_start,
                                                       uint32_t* __restrict__ items,
                                                       const uint32_t* __restrict__ key_offsets, int nscenes,
                                                       uint32_t* __restrict__ item_scene,
                                                       uint32_t* __restrict__ item_count) {
    __shared__ uint32_t ws[kScanWG / 64];
    uint32_t first = 0;
    if (RAW) {
        uint32_t part = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanWG) part += tile_start[t];
        block_exclusive_scan(part, ws, first);
    } else {
        first = tile_start[blockIdx.x];
    }
    const int64_t base = ((int64_t)blockIdx.x * kScanWG + threadIdx.x) * IPT;
    uint32_t keys[IPT];
    bool head[IPT];
    item_heads<IPT>(sk, n, base, item_max, keys, head);
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) c += head[j] ? 1u : 0u;
    uint32_t tot;
    uint32_t k = first + block_exclusive_scan(c, ws, tot);
    if (RAW && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *item_count = first + tot;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (!head[j]) continue;
        const int64_t i = base + j;
        const uint32_t key = keys[j];
        items[k] = (uint32_t)i;
        uint32_t koff = 0;
        if (item_scene) {
            int lo = 0, hi = nscenes - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (key_offsets[mid] <= key) lo = mid;
                else hi = mid - 1;
            }
            item_scene[k] = (uint32_t)lo;
            koff = key_offsets[lo];
        }
        items[n + k] = (key - koff) >> 6;
        ++k;
    }
}

template <int IPT>
static void launch_items_ipt(const uint32_t* sorted_keys, int64_t n, uint32_t* flags, uint32_t* scan_temp,
                             uint32_t* items, uint32_t* item_count, hipStream_t st, uint32_t item_max,
                             const uint32_t* key_offsets, int nscenes, uint32_t* item_scene, uint32_t* plan) {
    const int64_t tiles = (n + kScanWG * IPT - 1) / (kScanWG * IPT);
    // flags[0..tiles) holds the tile head counts (then, for many tiles,
    // their exclusive scan)
    hipLaunchKernelGGL(k_item_count<IPT>, dim3((unsigned)tiles), dim3(kScanWG), 0, st, sorted_keys, n, item_max,
                       flags, plan);
    if (tiles <= 2048) {
        hipLaunchKernelGGL((k_item_emit<true, IPT>), dim3((unsigned)tiles), dim3(kScanWG), 0, st, sorted_keys, n,
                           item_max, flags, items, key_offsets, nscenes, item_scene, item_count);
        return;
    }
    exclusive_scan(flags, tiles, scan_temp, item_count, st);
    hipLaunchKernelGGL((k_item_emit<false, IPT>), dim3((unsigned)tiles), dim3(kScanWG), 0, st, sorted_keys, n,
                       item_max, flags, items, key_offsets, nscenes, item_scene, item_count);
}

void launch_items(const uint32_t* sorted_keys, int64_t n, uint32_t* flags, uint32_t* scan_temp, uint32_t* items,
                  uint32_t* item_count, hipStream_t st, uint32_t item_max, const uint32_t* key_offsets,
                  int nscenes, uint32_t* item_scene, uint32_t* plan) {
    if (n <= 0) return;
    if (item_max == 0) {
        // a multiple of 64 up to kItemMax that still gives about four items
        // per compute unit: a small scene's few blocks are split so that
        // its P2G and G2P are not one item's serial chunks long
        const int64_t per = (n + 4 * 64 * (int64_t)device_cus() - 1) / (4 * 64 * (int64_t)device_cus());
        item_max = (uint32_t)(64 * (per < 1 ? 1 : (per > kItemMax / 64 ? kItemMax / 64 : per)));
    }
    // up to 2048 tiles the emit pass sums the tile counts itself (no scan
    // launches): 4 keys per thread up to 2M keys, then 16
    if (n <= (int64_t)2048 * kScanWG * 4)
        launch_items_ipt<4>(sorted_keys, n, flags, scan_temp, items, item_count, st, item_max, key_offsets, nscenes,
                            item_scene, plan);
    else
        launch_items_ipt<16>(sorted_keys, n, flags, scan_temp, items, item_count, st, item_max, key_offsets, nscenes,
                             item_scene, plan);
}

// ------------------------------------------------------- item decoding
struct ItemGeom {
    uint32_t s, e;
    bool escape;
    int ox, oy, oz;  // tile origin node
};
// key_offset: the scene's offset in the batch key space (0 for one scene).
__device__ __forceinline__ ItemGeom item_geom(const SceneDev& S, const ItemsDev& I, uint32_t it, uint32_t nitems,
                                              uint32_t key_offset) {
    ItemGeom g;
    g.s = I.start[it];
    g.e = (it + 1 < nitems) ? I.start[it + 1] : (uint32_t)I.n;
    const uint32_t blk = I.blk[it];
    (void)key_offset;
    g.escape = blk == (S.escape_key >> 6);
    uint32_t bx = blk % (uint32_t)S.kbx;
    uint32_t rest = blk / (uint32_t)S.kbx;
    uint32_t by = rest % (uint32_t)S.kby;
    uint32_t bz = rest / (uint32_t)S.kby;
    g.ox = 4 * (int)bx - 4;
    g.oy = 4 * (int)by - 4;
    g.oz = 4 * (int)bz - 4;
    return g;
}

// The scene an item belongs to: the kernel arguments for one scene, the
// batch table entry (read with scalar loads, the index is uniform) for a
// batch.  pbase is subtracted from the permutation's global slots.
template <bool BATCH>
__device__ __forceinline__ void item_scene(const SceneDev& S0, const ParticlesDev& P0, const ItemsDev& I,
                                           const SceneEntry* __restrict__ tab, uint32_t it, SceneDev& S,
                                           ParticlesDev& P, uint32_t& pbase, uint32_t& koff) {
    if constexpr (BATCH) {
        const uint32_t sc = I.scene[it];
        S = tab[sc].S;
        P = tab[sc].P;
        pbase = (uint32_t)tab[sc].pstart;
        koff = tab[sc].key_offset;
    } else {
        S = S0;
        P = P0;
        pbase = 0;
        koff = 0;
    }
}

__device__ __forceinline__ bool in_grid(const SceneDev& S, int i, int j, int k) {
    return i >= 0 && i < S.rx && j >= 0 && j < S.ry && k >= 0 && k < S.rz;
}

// ------------------------------------------------------------------ stress
// -V0 * P(Fe) Fe^T of one particle: plasticLame (src/material_point3d.cu:158-162)
// with fixed-corotated (:110-128) or Neo-Hookean (:131-155) stress, from an
// SVD Fe = U diag(s) V^T (U, V rotations).  The reference's expressions are
// evaluated in the principal frame, which avoids the cancellation in
// Fe - R (|Fe - R| ~ 1e-2 for snow):
//   fixed corotated: 2 mu (Fe - R) Fe^T + la (J-1) J I
//                  = U diag(2 mu (s_i - 1) s_i) U^T + la (J-1) J I
//   Neo-Hookean:     mu (Fe Fe^T - I) + la log(J) I
//                  = U diag(mu (s_i^2 - 1)) U^T + la log(J) I
// with J = s0 s1 s2 = det Fe.  The result is symmetric; tau holds xx, xy,
// xz, yy, yz, zz.  For Neo-Hookean the singular values are first raised to
// FLT_EPSILON and the projected Fe = U diag(s) V^T is returned in Fe (the
// reference writes it back into the particle, :142).
// Jp = det(Fp) (the hardening).
__device__ __forceinline__ void stress_from_svd(const float U[3][3], const float sv[3], const float V[3][3], float Jp,
                                                float mu0, float lam0, float V0, const MaterialDev& mat, float tau[6],
                                                float Fe[3][3]) {
    // exp to a few ulp (hardware exp2 of x log2 e); the reference's FP64
    // exp is narrowed to float (src/material_point3d.cu:160)
    const float factor = __expf(mat.hardening * (1.0f - Jp));
    const float mu = mu0 * factor, la = lam0 * factor;
    float s[3] = {sv[0], sv[1], sv[2]};
    float d[3], c;
    if (mat.model == MPM_MODEL_NEO_HOOKEAN) {
#pragma unroll
        for (int i = 0; i < 3; ++i) s[i] = fmaxf(s[i], MPM_FLT_EPS);
        float US[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) US[i][j] = U[i][j] * s[j];
        m3_mul_bt(US, V, Fe);
        const float J = s[0] * s[1] * s[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) d[i] = mu * (s[i] * s[i] - 1.0f);
        c = la * logf(J);
    } else {
        const float J = s[0] * s[1] * s[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) d[i] = (2.0f * mu) * ((s[i] - 1.0f) * s[i]);
        c = la * (J - 1.0f) * J;
    }
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int i = ii[k], j = jj[k];
        const float P = (U[i][0] * d[0] * U[j][0] + U[i][1] * d[1] * U[j][1]) + U[i][2] * d[2] * U[j][2] +
                        (i == j ? c : 0.0f);
        tau[k] = -(V0 * P);
    }
}

// Trial elastic deformation gradient Fn = (I + dt grad v) Fe
// (src/material_point3d.cu:57).
__device__ __forceinline__ void trial_deformation(const float Fe[3][3], const float grad[3][3], float dt,
                                                  float Fn[3][3]) {
    float L[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) L[i][j] = (i == j ? 1.0f : 0.0f) + dt * grad[i][j];
    m3_mul(L, Fe, Fn);
}

// Symmetric 3x3 matrices as (xx, xy, xz, yy, yz, zz).  sym_mul is the
// product of two symmetric matrices that commute (polynomials in one
// matrix), which is symmetric again.
__device__ __forceinline__ float sym_at(const float a[6], int i, int j) {
    const int k = i <= j ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5));
    return a[k];
}
__device__ __forceinline__ void sym_mul(const float a[6], const float b[6], float c[6]) {
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int i = ii[k], j = jj[k];
        c[k] = (sym_at(a, i, 0) * sym_at(b, 0, j) + sym_at(a, i, 1) * sym_at(b, 1, j)) + sym_at(a, i, 2) * sym_at(b, 2, j);
    }
}
// Leading principal minors of a symmetric matrix all positive (Sylvester):
// the matrix is positive definite.
__device__ __forceinline__ bool sym_positive_definite(const float a[6]) {
    const float m2 = a[0] * a[3] - a[1] * a[1];
    const float m3 = a[0] * (a[3] * a[5] - a[4] * a[4]) - a[1] * (a[1] * a[5] - a[4] * a[2]) +
                     a[2] * (a[1] * a[4] - a[3] * a[2]);
    return a[0] > 0.0f && m2 > 0.0f && m3 > 0.0f;
}

// The elastic case of updateDeformationGradient (src/material_point3d.cu:
// 53-70).  When every singular value of the trial Fn lies inside
// [1 - theta_c, 1 + theta_s] and det Fn > 0, the clamp changes nothing, so
// the reference's split gives Fe = Fn and leaves Fp as it is (exact
// arithmetic).  The singular values are the square roots of the eigenvalues
// of b = Fn Fn^T = I + E, so the test is that (hi^2 - 1) I - E and
// E + (1 - lo^2) I are positive definite.  The bounds are also capped at
// eigenvalues of E within +-0.05 so that the series below is accurate to
// ~3e-10 (materials with wider thresholds use the SVD path beyond that).
// Returns E (symmetric) and J = det Fn for the stress.
__device__ __forceinline__ bool elastic_trial(const float Fn[3][3], const MaterialDev& mat, float E[6], float& J) {
    const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int i = ii[k], j = jj[k];
        E[k] = (Fn[i][0] * Fn[j][0] + Fn[i][1] * Fn[j][1]) + Fn[i][2] * Fn[j][2] - (i == j ? 1.0f : 0.0f);
    }
    J = m3_det(Fn);
    const float hi = 1.0f + mat.crit_s, lo = 1.0f - mat.crit_c;
    const float ah = fminf(hi * hi - 1.0f, 0.05f);
    const float al = lo > 0.0f ? fminf(1.0f - lo * lo, 0.05f) : 0.05f;
    const float A[6] = {ah - E[0], -E[1], -E[2], ah - E[3], -E[4], ah - E[5]};
    const float C[6] = {E[0] + al, E[1], E[2], E[3] + al, E[4], E[5] + al};
    return J > 0.0f && sym_positive_definite(A) && sym_positive_definite(C);
}

// Fixed-corotated -V0 P(Fe) Fe^T (src/material_point3d.cu:110-128, with
// plasticLame :158-162) for Fe = Fn of an elastic step:
//   2 mu (Fe - R) Fe^T + la (J - 1) J I,  (Fe - R) Fe^T = b - sqrt(b)
// because R Fe^T = U diag(s) U^T = sqrt(Fe Fe^T) for Fe = U diag(s) V^T.
// With b = I + E, b - sqrt(b) is the series of 1 + x - sqrt(1 + x):
//   E/2 + E^2/8 - E^3/16 + 5 E^4/128 - 7 E^5/256 + O(E^6)
// (Horner form; the next term is below 21 |E|^6 / 1024 <= 3.2e-10).
__device__ __forceinline__ void elastic_fc_stress(const float E[6], float J, float Jp, float mu0, float lam0,
                                                  float V0, const MaterialDev& mat, float tau[6]) {
    const float factor = __expf(mat.hardening * (1.0f - Jp));
    const float mu = mu0 * factor, la = lam0 * factor;
    const bool dg[6] = {true, false, false, true, false, true};
    float M[6], T[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) M[k] = (dg[k] ? 5.0f / 128.0f : 0.0f) - (7.0f / 256.0f) * E[k];
    const float coef[3] = {-1.0f / 16.0f, 1.0f / 8.0f, 1.0f / 2.0f};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        sym_mul(E, M, T);
#pragma unroll
        for (int k = 0; k < 6; ++k) M[k] = T[k] + (dg[k] ? coef[r] : 0.0f);
    }
    sym_mul(E, M, T);  // b - sqrt(b)
    const float c = la * (J - 1.0f) * J;
#pragma unroll
    for (int k = 0; k < 6; ++k) tau[k] = -(V0 * ((2.0f * mu) * T[k] + (dg[k] ? c : 0.0f)));
}

// updateDeformationGradient, src/material_point3d.cu:53-70, general case:
// SVD of the trial Fn, clamp singular values to [1-theta_c, 1+theta_s],
// split into Fe and Fp.  Returns the SVD of the new Fe (U, clamped s, V)
// for the stress.  The reference's Fp = V diag(1/sc) U^T (Fn Fp) is
// evaluated as V diag(s/sc) V^T Fp, using U^T Fn = diag(s) V^T for the SVD
// Fn = U diag(s) V^T: the same matrix in exact arithmetic, two 3x3 products
// instead of four, and Fn Fp need not be kept through the SVD.
__device__ __forceinline__ void update_deformation(const float Fn[3][3], float Fe[3][3], float Fp[3][3],
                                                   const MaterialDev& mat, float U[3][3], float s[3],
                                                   float V[3][3]) {
    svd3(Fn, U, s, V);
    const float hi = 1.0f + mat.crit_s, lo = 1.0f - mat.crit_c;
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float sc = fmaxf(fminf(s[i], hi), lo);
        r[i] = s[i] * mpm_rcp(sc);
        s[i] = sc;
    }
    float US[3][3], VR[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            US[i][j] = U[i][j] * s[j];
            VR[i][j] = V[i][j] * r[j];
        }
    m3_mul_bt(US, V, Fe);
    float M[3][3], Fp_old[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Fp_old[i][j] = Fp[i][j];
    m3_mul_bt(VR, V, M);
    m3_mul(M, Fp_old, Fp);
}

__device__ __forceinline__ void load_m3(const float* base, int64_t N, uint32_t p, float M[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[i][j] = base[(int64_t)(3 * i + j) * N + p];
}
__device__ __forceinline__ void store_m3(float* base, int64_t N, uint32_t p, const float M[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) base[(int64_t)(3 * i + j) * N + p] = M[i][j];
}

// -------------------------------------------------------------------- P2G
// particlesToGrid: src/mpm_solver3d.cu:237-294 (PIC-FLIP),
// src/apic_mpm_solver3d.cu:245-304 (APIC); MODE 1 is initialize()'s
// P2GMass (src/mpm_solver3d.cu:136-171).
//
// One wave per work item.  Lanes hold consecutive sorted particles, so the
// particles of one cell sit in consecutive lanes and share all W^3 target
// nodes.  Each contribution is first summed over the lanes of its cell (a
// segmented scan in DPP, below), then only the last lane of each cell adds
// it to the LDS tile with a plain read-add-write: within one instruction the
// writing lanes belong to different cells, so they hit different nodes, and
// the single wave orders successive writes.  This avoids LDS float atomics.

// Segmented inclusive scan toward the LAST lane of each segment, entirely
// in DPP (VALU, no LDS traffic): Hillis-Steele within each 16-lane row with
// row_shr:1,2,4,8, then row_bcast:15 (last lane of rows 0 and 2 into rows 1
// and 3) and row_bcast:31 (lane 31 into rows 2 and 3), as in a wave64 scan.
// Each step is one v_fmac_f32 with a DPP source per value,
//   v += shifted(v) * flag,
// where the per-lane flag is 1.0 when the source lane lies in the same
// segment and 0.0 otherwise (bound_ctrl gives 0 for sources outside the
// row; row_mask keeps the rows a broadcast does not target).  The compiler
// does not fold a DPP move into a tied FMAC, so the step is written as
// inline assembly; each block starts with s_nop 1, the 2 wait states a DPP
// read needs after a VALU write of its source (the values are computed just
// before the first block).  The 5 wait states a DPP needs after a VALU
// write of EXEC do not arise: these kernels change EXEC with scalar
// instructions only (no v_cmpx).  Within a block the 4 or 7 values are
// independent, so successive blocks need no further waits.  Steps that no
// lane needs are skipped (uniform branches).
struct SegPlan {
    float f1, f2, f4, f8, f15, f31;  // per-lane multipliers of each step
    bool s1, s2, s4, s8, s15, s31;   // uniform: step needed
    bool writer;                     // last active lane of its segment
};
// headmask: lanes that start a segment (lane 0 always does).
__device__ __forceinline__ SegPlan seg_plan(uint64_t headmask, bool active, int lane) {
    SegPlan sp;
    const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const int start = 63 - __clzll(headmask & le);
    const int rowstart = lane & ~15;
    sp.f1 = start <= lane - 1 ? 1.0f : 0.0f;
    sp.f2 = start <= lane - 2 ? 1.0f : 0.0f;
    sp.f4 = start <= lane - 4 ? 1.0f : 0.0f;
    sp.f8 = start <= lane - 8 ? 1.0f : 0.0f;
    const bool c15 = ((lane >> 4) & 1) && start < rowstart;
    const bool c31 = lane >= 32 && start <= 31;
    sp.f15 = c15 ? 1.0f : 0.0f;
    sp.f31 = c31 ? 1.0f : 0.0f;
    const int run = lane - (start > rowstart ? start : rowstart);
    sp.s1 = __ballot(run >= 1) != 0ull;
    sp.s2 = __ballot(run >= 2) != 0ull;
    sp.s4 = __ballot(run >= 4) != 0ull;
    sp.s8 = __ballot(run >= 8) != 0ull;
    sp.s15 = __ballot(c15) != 0ull;
    sp.s31 = __ballot(c31) != 0ull;
    sp.writer = active && (lane == 63 || ((headmask >> (lane + 1)) & 1ull));
    return sp;
}

#define MPM_FMAC_DPP(CTRL) "v_fmac_f32_dpp %0, %0, %1 " CTRL " bank_mask:0xf bound_ctrl:1"
#define MPM_FMAC_DPP7(CTRL)                                                                    \
    "s_nop 1\n\t"                                                                              \
    "v_fmac_f32_dpp %0, %0, %7 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %1, %1, %7 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %2, %2, %7 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %3, %3, %7 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %4, %4, %7 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %5, %5, %7 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %6, %6, %7 " CTRL " bank_mask:0xf bound_ctrl:1"

#define MPM_FMAC_DPP4(CTRL)                                                                    \
    "s_nop 1\n\t"                                                                              \
    "v_fmac_f32_dpp %0, %0, %4 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %1, %1, %4 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %2, %2, %4 " CTRL " bank_mask:0xf bound_ctrl:1\n\t"                       \
    "v_fmac_f32_dpp %3, %3, %4 " CTRL " bank_mask:0xf bound_ctrl:1"

#define MPM_SCAN_STEP(NV, v, CTRL, flag)                                                          \
    do {                                                                                        \
        if constexpr (NV == 4) {                                                                \
            asm volatile(MPM_FMAC_DPP4(CTRL) : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]) : "v"(flag)); \
        } else if constexpr (NV == 7) {                                                         \
            asm volatile(MPM_FMAC_DPP7(CTRL)                                                    \
                         : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), \
                           "+v"(v[6])                                                           \
                         : "v"(flag));                                                          \
        } else {                                                                                \
            for (int f_ = 0; f_ < NV; ++f_)                                                     \
                asm volatile("s_nop 1\n\t" MPM_FMAC_DPP(CTRL) : "+v"(v[f_]) : "v"(flag));        \
        }                                                                                       \
    } while (0)

// Segmented scan of the first NV values of v.
template <int NV, int N>
__device__ __forceinline__ void seg_scan(float (&v)[N], const SegPlan& sp) {
    static_assert(NV == 7 || NV == 4 || NV == 1, "scan blocks are written for 7, 4 or 1 values");
    static_assert(NV <= N, "fewer values than the scan covers");
    if (sp.s1) MPM_SCAN_STEP(NV, v, "row_shr:1 row_mask:0xf", sp.f1);
    if (sp.s2) MPM_SCAN_STEP(NV, v, "row_shr:2 row_mask:0xf", sp.f2);
    if (sp.s4) MPM_SCAN_STEP(NV, v, "row_shr:4 row_mask:0xf", sp.f4);
    if (sp.s8) MPM_SCAN_STEP(NV, v, "row_shr:8 row_mask:0xf", sp.f8);
    if (sp.s15) MPM_SCAN_STEP(NV, v, "row_bcast:15 row_mask:0xa", sp.f15);
    if (sp.s31) MPM_SCAN_STEP(NV, v, "row_bcast:31 row_mask:0xc", sp.f31);
}

template <int INTERP, int APIC, int MODE, bool BATCH>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(P2GWaves<INTERP, MODE>::value))) k_p2g(SceneDev S0, ParticlesDev P0, ItemsDev I,
                                            const SceneEntry* __restrict__ tab) {
    constexpr int W = Stencil<INTERP>::W;
    constexpr int T = 4 + W - 1;
    constexpr int TN = T * T * T;
    constexpr int NF = MODE == 0 ? 7 : (MODE == 1 ? 1 : 4);
    static_assert(MODE != 2 || APIC, "the folded grid serves APIC only");
    __shared__ float tile[NF * TN];
    const uint32_t nitems = *I.count;
    const int lane = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        // a fused batch step: scenes whose last step already scattered this P2G
        if (BATCH && tab[I.scene[it]].prescat != 0u) continue;
        SceneDev S;
        ParticlesDev P;
        uint32_t pbase, koff;
        item_scene<BATCH>(S0, P0, I, tab, it, S, P, pbase, koff);
        const int64_t N = P.n;
        ItemGeom g = item_geom(S, I, it, nitems, koff);
        if (g.escape) {  // no stencil node inside the grid
            if (I.snap_k)
                for (uint32_t i = g.s + lane; i < g.e; i += 64) {
                    I.snap_k[i] = I.sorted_keys[i];
                    I.snap_p[i] = I.perm[i];
                }
            continue;
        }
        for (int q = lane; q < NF * TN; q += 64) tile[q] = 0.0f;
        __syncthreads();
        for (uint32_t base = g.s; base < g.e; base += 64) {
            const uint32_t i = base + lane;
            const bool active = i < g.e;
            const uint32_t ic = active ? i : g.s;
            // segment structure of this chunk (inactive lanes form their own tail segment)
            const uint32_t key = active ? I.sorted_keys[ic] : 0xFFFFFFFFu;
            const uint32_t prev = __shfl_up(key, 1, 64);
            const SegPlan sp = seg_plan(__ballot(lane == 0 || prev != key), active, lane);
            const bool writer = sp.writer;
            const uint32_t praw = I.perm[ic];
            const uint32_t p = praw - pbase;
            if (I.snap_k && active) {  // the sorted order, kept for the next step's incremental sort
                I.snap_k[i] = key;
                I.snap_p[i] = praw;
            }
            const float x0 = P.x[p], x1 = P.x[N + p], x2 = P.x[2 * N + p];
            int bx, by, bz;
            float Nx[W], Ny[W], Nz[W], dNx[W], dNy[W], dNz[W];
            axis_weights<INTERP>(x0, S.ox, S.h, S.inv_h, bx, Nx, dNx);
            axis_weights<INTERP>(x1, S.oy, S.h, S.inv_h, by, Ny, dNy);
            axis_weights<INTERP>(x2, S.oz, S.h, S.inv_h, bz, Nz, dNz);
            const int lx = bx - g.ox, ly = by - g.oy, lz = bz - g.oz;
            // tile cell of the stencil's first node; node (a, b, c) is at
            // tq[(c * T + b) * T + a] (constant offsets)
            float* const tq = tile + ((lz * T + ly) * T + lx);
            const float m = active ? P.m[p] : 0.0f;
            if (MODE == 1) {
#pragma unroll
                for (int a = 0; a < W; ++a)
#pragma unroll
                    for (int b = 0; b < W; ++b)
#pragma unroll
                        for (int c = 0; c < W; ++c) {
                            float v[1] = {m * ((Nx[a] * Ny[b]) * Nz[c])};
                            seg_scan<1>(v, sp);
                            const int q = (c * T + b) * T + a;
                            if (writer) tq[q] += v[0];
                        }
                continue;
            }
            // cached -V0 P Fe^T (symmetric), written by the previous G2P or
            // k_stress; scaled by dt for the folded grid (MODE 2), whose
            // planes 1..3 receive p + dt f
            float vs[3][3];
            {
                const float act = active ? (MODE == 2 ? S.dt : 1.0f) : 0.0f;
                const float t0 = act * P.tau[p], t1 = act * P.tau[N + p], t2 = act * P.tau[2 * N + p];
                const float t3 = act * P.tau[3 * N + p], t4 = act * P.tau[4 * N + p], t5 = act * P.tau[5 * N + p];
                vs[0][0] = t0; vs[0][1] = t1; vs[0][2] = t2;
                vs[1][0] = t1; vs[1][1] = t3; vs[1][2] = t4;
                vs[2][0] = t2; vs[2][1] = t4; vs[2][2] = t5;
            }
            const float v0 = P.v[p], v1 = P.v[N + p], v2 = P.v[2 * N + p];
            // Tensor-product factorisation of the node contributions
            // (w = Nx[a] Ny[b] Nz[c]):
            //   mass     m w                 = (m Nx[a]) yz
            //   momentum m w (v + C (x_i - x_p)) = mip (cx(a) + cyz(b,c))
            //   force    vs grad w           = F0(a) yz + F1(a) dyz + F2(a) ydz
            float dxn[3][W];
            float Cm[3][3];
            if (APIC) {
                float Bm[3][3];
                load_m3(P.B, N, p, Bm);
#pragma unroll
                for (int i2 = 0; i2 < 3; ++i2)
#pragma unroll
                    for (int j2 = 0; j2 < 3; ++j2) Cm[i2][j2] = Bm[i2][j2] * S.dinv;
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    dxn[0][j] = (S.ox + S.h * (float)(bx + j)) - x0;
                    dxn[1][j] = (S.oy + S.h * (float)(by + j)) - x1;
                    dxn[2][j] = (S.oz + S.h * (float)(bz + j)) - x2;
                }
            }
            if constexpr (MODE == 2) {
                // the folded grid: momentum plus force of node (a, b, c) as
                // yz R(a) + Nx[a] Syz(b,c) (the regrouping of the fused
                // kernel's P2G half), (y, z) pairs outermost
                const float vk3[3] = {v0, v1, v2};
                float ma[W], R[W][3];
#pragma unroll
                for (int a = 0; a < W; ++a) {
                    ma[a] = m * Nx[a];
#pragma unroll
                    for (int k = 0; k < 3; ++k) R[a][k] = fmaf(ma[a], vk3[k] + Cm[k][0] * dxn[0][a], vs[k][0] * dNx[a]);
                }
#pragma unroll
                for (int b = 0; b < W; ++b)
#pragma unroll
                    for (int c = 0; c < W; ++c) {
                        const float yzc = Ny[b] * Nz[c], dyzc = dNy[b] * Nz[c], ydzc = Ny[b] * dNz[c];
                        const float myz = m * yzc;
                        float Syz[3];
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                            Syz[k] = fmaf(myz, Cm[k][1] * dxn[1][b] + Cm[k][2] * dxn[2][c],
                                          vs[k][1] * dyzc + vs[k][2] * ydzc);
#pragma unroll
                        for (int a = 0; a < W; ++a) {
                            float v[NF];
                            v[0] = ma[a] * yzc;
#pragma unroll
                            for (int k = 0; k < 3; ++k) v[1 + k] = fmaf(yzc, R[a][k], Nx[a] * Syz[k]);
                            seg_scan<NF>(v, sp);
                            const int q = (c * T + b) * T + a;
                            if (writer) {
                                float o[NF];
#pragma unroll
                                for (int f = 0; f < NF; ++f) o[f] = tq[f * TN + q];
#pragma unroll
                                for (int f = 0; f < NF; ++f) tq[f * TN + q] = o[f] + v[f];
                            }
                        }
                    }
                continue;
            }
            float yz[W][W], dyz[W][W], ydz[W][W];
#pragma unroll
            for (int b = 0; b < W; ++b)
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    yz[b][c] = Ny[b] * Nz[c];
                    dyz[b][c] = dNy[b] * Nz[c];
                    ydz[b][c] = Ny[b] * dNz[c];
                }
#pragma unroll
            for (int a = 0; a < W; ++a) {
                const float ma = m * Nx[a];
                float F0[3], F1[3], F2[3], cx[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    F0[k] = vs[k][0] * dNx[a];
                    F1[k] = vs[k][1] * Nx[a];
                    F2[k] = vs[k][2] * Nx[a];
                    const float vk = k == 0 ? v0 : (k == 1 ? v1 : v2);
                    cx[k] = APIC ? vk + Cm[k][0] * dxn[0][a] : vk;
                }
#pragma unroll
                for (int b = 0; b < W; ++b)
#pragma unroll
                    for (int c = 0; c < W; ++c) {
                        const float mip = ma * yz[b][c];
                        float v[7];
                        v[0] = mip;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const float ck = APIC ? cx[k] + (Cm[k][1] * dxn[1][b] + Cm[k][2] * dxn[2][c]) : cx[k];
                            const float fk = (F0[k] * yz[b][c] + F1[k] * dyz[b][c]) + F2[k] * ydz[b][c];
                            if (MODE == 2) {
                                v[1 + k] = fmaf(mip, ck, fk);
                            } else {
                                v[1 + k] = mip * ck;
                                v[4 + k] = fk;
                            }
                        }
                        seg_scan<NF>(v, sp);
                        const int q = (c * T + b) * T + a;
                        if (writer) {
                            // all reads before the writes: one LDS round trip per node
                            float o[NF];
#pragma unroll
                            for (int f = 0; f < NF; ++f) o[f] = tq[f * TN + q];
#pragma unroll
                            for (int f = 0; f < NF; ++f) tq[f * TN + q] = o[f] + v[f];
                        }
                    }
            }
        }
        __syncthreads();
        for (int q = lane; q < TN; q += 64) {
            const int lx = q % T, ly = (q / T) % T, lz = q / (T * T);
            const int gi = g.ox + lx, gj = g.oy + ly, gk = g.oz + lz;
            if (!in_grid(S, gi, gj, gk)) continue;
            const int64_t ni = node_index(S, gi, gj, gk);
            bool any = false;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const float val = tile[f * TN + q];
                if (val != 0.0f) {
                    unsafeAtomicAdd(&S.grid[(int64_t)f * S.plane + ni], val);
                    any = true;
                }
            }
            if (any) S.touched[storage_block(S, gi, gj, gk)] = 1u;
        }
        __syncthreads();
    }
}

template <int INTERP, int APIC>
static void p2g_dispatch_mode(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, const SceneEntry* tab,
                              int mode, int wgs, hipStream_t st) {
    if constexpr (APIC) {
        if (mode == 2) {
            if (tab) MPM_PERSISTENT((k_p2g<INTERP, APIC, 2, true>), 64, wgs, st, S, P, I, tab);
            else MPM_PERSISTENT((k_p2g<INTERP, APIC, 2, false>), 64, wgs, st, S, P, I, tab);
            return;
        }
    }
    if (tab) MPM_PERSISTENT((k_p2g<INTERP, APIC, 0, true>), 64, wgs, st, S, P, I, tab);
    else if (mode == 0) MPM_PERSISTENT((k_p2g<INTERP, APIC, 0, false>), 64, wgs, st, S, P, I, tab);
    else MPM_PERSISTENT((k_p2g<INTERP, APIC, 1, false>), 64, wgs, st, S, P, I, tab);
}
template <int INTERP>
static void p2g_dispatch_apic(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, const SceneEntry* tab,
                              int apic, int mode, int wgs, hipStream_t st) {
    if (apic) p2g_dispatch_mode<INTERP, 1>(S, P, I, tab, mode, wgs, st);
    else p2g_dispatch_mode<INTERP, 0>(S, P, I, tab, mode, wgs, st);
}
void launch_p2g(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, int interp, int apic, int mode,
                int grid_wgs, hipStream_t st, const SceneEntry* tab) {
    if (I.n <= 0) return;
    switch (interp) {
        case MPM_INTERP_LINEAR: p2g_dispatch_apic<MPM_INTERP_LINEAR>(S, P, I, tab, apic, mode, grid_wgs, st); break;
        case MPM_INTERP_QUADRATIC_BSPLINE:
            p2g_dispatch_apic<MPM_INTERP_QUADRATIC_BSPLINE>(S, P, I, tab, apic, mode, grid_wgs, st);
            break;
        default: p2g_dispatch_apic<MPM_INTERP_CUBIC_BSPLINE>(S, P, I, tab, apic, mode, grid_wgs, st); break;
    }
}

// A particle whose cell key changed in this step goes on the move list of
// the incremental sort: the wave takes sub-list (global wave index mod
// kMoveLists), one counter atomic per wave, then each moved lane writes
// (new key, slot, old sorted position) at its rank.  Past the sub-list's
// capacity nothing is stored and the overflow word sends the next step to
// the full sort.  In two halves, so that a kernel can issue the atomic
// early and use its result late (the wave need not wait for it).
struct MoveTicket {
    uint64_t mask;  // the wave's moved lanes
    uint32_t base;  // the leader lane's atomic result
    uint32_t list;  // sub-list index
};
__device__ __forceinline__ MoveTicket move_begin(const ItemsDev& I, bool moved) {
    MoveTicket t;
    t.mask = __ballot(moved);
    t.base = 0;
    t.list = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (uint32_t)(kMoveLists - 1);
    if (t.mask == 0ull) return t;
    const int lane = threadIdx.x & 63;
    if (lane == __ffsll((unsigned long long)t.mask) - 1)
        t.base = atomicAdd(&I.plan[kPlanLists + t.list * kMoveCountStride], (uint32_t)__popcll(t.mask));
    return t;
}
__device__ __forceinline__ void move_end(const ItemsDev& I, const MoveTicket& t, uint32_t key, uint32_t slot,
                                         uint32_t pos) {
    if (t.mask == 0ull) return;
    const int lane = threadIdx.x & 63;
    const uint32_t base = __shfl(t.base, __ffsll((unsigned long long)t.mask) - 1, 64);
    if ((t.mask >> lane) & 1ull) {
        const uint32_t idx = base + (uint32_t)__popcll(t.mask & ((1ull << lane) - 1ull));
        if (idx < I.move_cap) {
            const size_t e = (size_t)t.list * I.move_cap + idx;
            I.moves->key[e] = key;
            I.moves->slot[e] = slot;
            I.moves->pos[e] = pos;
        } else {
            I.plan[kPlanOverflow] = 1u;
        }
    }
}

// ------------------------------------------------------ XCD-aware items
// Workgroups are dispatched to the 8 XCDs in turn (workgroup b on XCD b % 8),
// and each XCD has its own L2.  Consecutive work items are neighbouring key
// blocks whose tiles share halo nodes, so a persistent kernel hands out
// chunks of kXcdChunk consecutive items per XCD: XCD x takes chunks x,
// x + 8, ..., and its workgroups step through the items of those chunks.
// Which workgroup runs on which XCD only changes the L2 hits, not the
// coverage: every item is taken exactly once (fewer than 8 workgroups walk
// the plain order).
#ifndef MPM_XCD_CHUNK
#define MPM_XCD_CHUNK 16
#endif
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdChunk = MPM_XCD_CHUNK;
struct XcdWalk {
    uint32_t x, step, k;
    bool on;
};
__device__ __forceinline__ XcdWalk xcd_walk() {
    XcdWalk w;
    w.on = kXcdChunk > 0 && gridDim.x >= kXcds;
    w.x = blockIdx.x % kXcds;
    w.step = w.on ? (gridDim.x - 1 - w.x) / kXcds + 1 : gridDim.x;
    w.k = w.on ? blockIdx.x / kXcds : blockIdx.x;
    return w;
}
// Item at the walk's k-th position (increasing in k for one XCD).
__device__ __forceinline__ uint32_t xcd_item(const XcdWalk& w, uint32_t k) {
    if constexpr (kXcdChunk == 0) {
        return k;
    } else {
        if (!w.on) return k;
        return (w.x + kXcds * (k / kXcdChunk)) * kXcdChunk + k % kXcdChunk;
    }
}

// -------------------------------------------------------------------- G2P
// Folded-grid tile of one item, in two halves.  fetch_folded_tile issues
// the lane's tile-node loads (mass and p + dt f) as global_load_lds_dword:
// the data goes straight from HBM into the staging area in LDS (word
// f * TN + q for value f of tile node q: 4 TN words, the wave's 64
// consecutive nodes land at one wave-uniform base) without passing through
// VGPRs.
// commit_folded_tile waits for them (vmcnt(0): the compiler does not track
// these LDS writes) and turns them into v* (the fused grid update) in the
// tile.  The G2P loop fetches the next item's tile before the current item's
// particles, so the loads are in flight during that work.  Nodes outside the
// grid issue no load; commit writes 0 for them without reading the stage.
typedef __attribute__((address_space(3))) void lds_void;
template <int T, int TR, int WG>
__device__ __forceinline__ void fetch_folded_tile(const SceneDev& S, const ItemGeom& g, float* stage) {
    constexpr int TN = T * T * T;
#pragma unroll 1  // unrolled, the address arithmetic of all rounds stays live (spills at 64 threads)
    for (int r = 0; r < TR; ++r) {
        const int q = (int)threadIdx.x + r * WG;
        // the wave's first node this round: its lanes write words base + lane
        float* const wst = stage + r * WG + (int)(threadIdx.x & ~63u);
        const int lx = q % T, ly = (q / T) % T, lz = q / (T * T);
        const int gi = g.ox + lx, gj = g.oy + ly, gk = g.oz + lz;
        if (q < TN && !g.escape && in_grid(S, gi, gj, gk)) {
            const int64_t ni = node_index(S, gi, gj, gk);
            const int64_t pl = S.plane;
            __builtin_amdgcn_global_load_lds((const void*)(S.grid + ni), (lds_void*)wst, 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(S.grid + (GF_VEL + 0) * pl + ni), (lds_void*)(wst + TN), 4,
                                             0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(S.grid + (GF_VEL + 1) * pl + ni),
                                             (lds_void*)(wst + 2 * TN), 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(S.grid + (GF_VEL + 2) * pl + ni),
                                             (lds_void*)(wst + 3 * TN), 4, 0, 0);
        }
    }
}
template <int T, int TR, int WG>
__device__ __forceinline__ void commit_folded_tile(const SceneDev& S, const ItemGeom& g, const float* stage,
                                                   float* tile) {
    constexpr int TN = T * T * T;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll 1  // as in fetch_folded_tile (the cubic stencil spills at 64 threads otherwise)
    for (int r = 0; r < TR; ++r) {
        const int q = (int)threadIdx.x + r * WG;
        if (q >= TN) break;
        const int lx = q % T, ly = (q / T) % T, lz = q / (T * T);
        const int gi = g.ox + lx, gj = g.oy + ly, gk = g.oz + lz;
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
        if (in_grid(S, gi, gj, gk))
            grid_node_update_folded(S, gi, gj, gk, stage[q], stage[TN + q], stage[2 * TN + q], stage[3 * TN + q], v0,
                                    v1, v2);
        tile[q] = v0;
        tile[TN + q] = v1;
        tile[2 * TN + q] = v2;
    }
}

// -------------------------------------------------------------------- G2P
// gridToParticles: src/mpm_solver3d.cu:316-379 (PIC-FLIP),
// src/apic_mpm_solver3d.cu:326-388 (APIC), with the optional
// particlesCollision (:428-445) and advectParticles (:381-391) fused in.
// SRC 2 is initialize()'s G2PVolume (src/mpm_solver3d.cu:173-214).
template <int INTERP, int APIC, int SRC, bool BATCH, int WG>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(G2PWaves<INTERP, APIC, WG>::value))) k_g2p(SceneDev S0, ParticlesDev P0, ItemsDev I, const SceneEntry* __restrict__ tab,
                                            int advect, int pcollide, uint32_t* __restrict__ next_keys) {
    constexpr int W = Stencil<INTERP>::W;
    constexpr int T = 4 + W - 1;
    constexpr int TN = T * T * T;
    constexpr int NF = SRC == 2 ? 1 : (APIC ? 3 : 6);
    static_assert(SRC != 3 || APIC, "the folded grid serves APIC only");
    // folded grid: the next item's tile loads overlap this item
    constexpr bool PREFETCH = SRC == 3;
    constexpr int TR = PREFETCH ? (TN + WG - 1) / WG : 1;
    // and each particle's Fe, Jp, mu0, lam0, V0 go straight into LDS when
    // the particle is picked, to be read after the gather (not for cubic:
    // its larger tile would cut the workgroups per CU); Fp is read only by
    // a wave with a plastic update
    constexpr bool PSTAGE = PREFETCH && INTERP != MPM_INTERP_CUBIC_BSPLINE;
    constexpr int PW = 13;  // staged words per particle
    __shared__ float tile[NF * TN];
    __shared__ float stage[PREFETCH ? 4 * TN : 1];
    __shared__ float pstage[PSTAGE ? PW * WG : 1];
    const uint32_t nitems = *I.count;
    ItemGeom gnext{};
    const XcdWalk walk = xcd_walk();  // items in XCD-local chunks (neighbouring tiles share an L2)
    const uint32_t it0 = xcd_item(walk, walk.k);
    if constexpr (PREFETCH) {
        if (it0 < nitems) {
            SceneDev Sn;
            ParticlesDev Pn;
            uint32_t pbn, kon;
            item_scene<BATCH>(S0, P0, I, tab, it0, Sn, Pn, pbn, kon);
            gnext = item_geom(Sn, I, it0, nitems, kon);
            fetch_folded_tile<T, TR, WG>(Sn, gnext, stage);
        }
    }
    for (uint32_t k = walk.k, it = it0; it < nitems; k += walk.step, it = xcd_item(walk, k)) {
        SceneDev S;
        ParticlesDev P;
        uint32_t pbase, koff;
        item_scene<BATCH>(S0, P0, I, tab, it, S, P, pbase, koff);
        const int64_t N = P.n;
        ItemGeom g = PREFETCH ? gnext : item_geom(S, I, it, nitems, koff);
        if (PREFETCH) {
            if (!g.escape) commit_folded_tile<T, TR, WG>(S, g, stage, tile);
        } else if (!g.escape) {
            for (int q = threadIdx.x; q < TN; q += WG) {
                const int lx = q % T, ly = (q / T) % T, lz = q / (T * T);
                const int gi = g.ox + lx, gj = g.oy + ly, gk = g.oz + lz;
                float vals[NF];
#pragma unroll
                for (int f = 0; f < NF; ++f) vals[f] = 0.0f;
                if (in_grid(S, gi, gj, gk)) {
                    const int64_t ni = node_index(S, gi, gj, gk);
                    const float* G = S.grid;
                    const int64_t pl = S.plane;
                    if (SRC == 2) {
                        vals[0] = G[ni];
                    } else if (SRC == 1) {
                        vals[0] = G[(GF_VSTAR + 0) * pl + ni];
                        vals[1] = G[(GF_VSTAR + 1) * pl + ni];
                        vals[2] = G[(GF_VSTAR + 2) * pl + ni];
                        if (!APIC) {  // FLIP increment v* - v
                            vals[3] = vals[0] - G[(GF_VEL + 0) * pl + ni];
                            vals[4] = vals[1] - G[(GF_VEL + 1) * pl + ni];
                            vals[5] = vals[2] - G[(GF_VEL + 2) * pl + ni];
                        }
                    } else if (SRC == 3) {
                        grid_node_update_folded(S, gi, gj, gk, G[ni], G[(GF_VEL + 0) * pl + ni],
                                                G[(GF_VEL + 1) * pl + ni], G[(GF_VEL + 2) * pl + ni], vals[0],
                                                vals[1], vals[2]);
                    } else {
                        float vx, vy, vz, wx, wy, wz;
                        grid_node_update(S, gi, gj, gk, G[ni], G[(GF_VEL + 0) * pl + ni], G[(GF_VEL + 1) * pl + ni],
                                         G[(GF_VEL + 2) * pl + ni], G[(GF_FORCE + 0) * pl + ni],
                                         G[(GF_FORCE + 1) * pl + ni], G[(GF_FORCE + 2) * pl + ni], vx, vy, vz, wx, wy,
                                         wz);
                        vals[0] = wx; vals[1] = wy; vals[2] = wz;
                        if (!APIC) { vals[3] = wx - vx; vals[4] = wy - vy; vals[5] = wz - vz; }
                    }
                }
#pragma unroll
                for (int f = 0; f < NF; ++f) tile[f * TN + q] = vals[f];
            }
        }
        __syncthreads();
        if constexpr (PREFETCH) {
            const uint32_t itn = xcd_item(walk, k + walk.step);
            if (itn < nitems) {
                SceneDev Sn;  // the next item's scene (a batch's items span scenes)
                ParticlesDev Pn;
                uint32_t pbn, kon;
                item_scene<BATCH>(S0, P0, I, tab, itn, Sn, Pn, pbn, kon);
                gnext = item_geom(Sn, I, itn, nitems, kon);
                fetch_folded_tile<T, TR, WG>(Sn, gnext, stage);
            }
        }
        for (uint32_t i = g.s + threadIdx.x; i < g.e; i += WG) {
            const uint32_t p = I.perm[i] - pbase;
            // the key it was sorted by, for the move list: loaded with the
            // slot, long before the comparison needs it
            const uint32_t kold = I.plan ? I.sorted_keys[i] : 0u;
            uint32_t knew = 0;
            MoveTicket ticket{};
            float* const pst = pstage + (threadIdx.x >> 6) * (PW * 64);  // the wave's region
            if constexpr (PSTAGE) {
#pragma unroll 1
                for (int f = 0; f < 9; ++f)
                    __builtin_amdgcn_global_load_lds((const void*)(P.Fe + (int64_t)f * N + p), (lds_void*)(pst + f * 64),
                                                     4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.Jp + p), (lds_void*)(pst + 9 * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.mu0 + p), (lds_void*)(pst + 10 * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.lam0 + p), (lds_void*)(pst + 11 * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.V0 + p), (lds_void*)(pst + 12 * 64), 4, 0, 0);
            }
            const float x0 = P.x[p], x1 = P.x[N + p], x2 = P.x[2 * N + p];
            int bx = 0, by = 0, bz = 0;
            float Nx[W], Ny[W], Nz[W], dNx[W], dNy[W], dNz[W];
            if (!g.escape) {
                axis_weights<INTERP>(x0, S.ox, S.h, S.inv_h, bx, Nx, dNx);
                axis_weights<INTERP>(x1, S.oy, S.h, S.inv_h, by, Ny, dNy);
                axis_weights<INTERP>(x2, S.oz, S.h, S.inv_h, bz, Nz, dNz);
            }
            const int lx = bx - g.ox, ly = by - g.oy, lz = bz - g.oz;
            if (SRC == 2) {
                float density = 0.0f;
                if (!g.escape) {
#pragma unroll
                    for (int a = 0; a < W; ++a)
#pragma unroll
                        for (int b = 0; b < W; ++b)
#pragma unroll
                            for (int c = 0; c < W; ++c) {
                                const float w = (Nx[a] * Ny[b]) * Nz[c];
                                const float mi = tile[((lz + c) * T + (ly + b)) * T + (lx + a)];
                                density += (mi * w) * S.inv_cell_vol;
                            }
                }
                P.V0[p] = P.m[p] / density;
                continue;
            }
            const float vp0 = APIC ? 0.0f : P.v[p];
            const float vp1 = APIC ? 0.0f : P.v[N + p];
            const float vp2 = APIC ? 0.0f : P.v[2 * N + p];
            float vpic0 = 0.0f, vpic1 = 0.0f, vpic2 = 0.0f;
            float vf0 = vp0, vf1 = vp1, vf2 = vp2;
            float grad[3][3], Bm[3][3];
#pragma unroll
            for (int i2 = 0; i2 < 3; ++i2)
#pragma unroll
                for (int j2 = 0; j2 < 3; ++j2) { grad[i2][j2] = 0.0f; Bm[i2][j2] = 0.0f; }
            if (!g.escape) {
                float dxn[3][W];
                if (APIC) {
#pragma unroll
                    for (int j = 0; j < W; ++j) {
                        dxn[0][j] = (S.ox + S.h * (float)(bx + j)) - x0;
                        dxn[1][j] = (S.oy + S.h * (float)(by + j)) - x1;
                        dxn[2][j] = (S.oz + S.h * (float)(bz + j)) - x2;
                    }
                }
                // Nested tensor-product sums, z innermost: with w = Nx Ny Nz,
                // grad w = (dNx Ny Nz, Nx dNy Nz, Nx Ny dNz) / h and the APIC
                // moment w (x_i - x_p), every sum over the stencil is
                //   sum_a X[a] sum_b Y[b] sum_c Z[c] v(a, b, c)
                // for one of a few per-axis weight vectors; the z sums (3
                // weights x 3 components per node) feed the y sums (5 per
                // (a, b)), which feed the x sums (7 per a).
                float zdx[W], ydx[W];
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    zdx[j] = APIC ? Nz[j] * dxn[2][j] : 0.0f;
                    ydx[j] = APIC ? Ny[j] * dxn[1][j] : 0.0f;
                }
                // tile cell of the stencil's first node (constant offsets per node)
                const float* const tq = tile + ((lz * T + ly) * T + lx);
#pragma unroll
                for (int a = 0; a < W; ++a) {
                    float Y0[3] = {0.0f, 0.0f, 0.0f}, Yd[3] = {0.0f, 0.0f, 0.0f}, Yz[3] = {0.0f, 0.0f, 0.0f};
                    float Yx[3] = {0.0f, 0.0f, 0.0f}, Yzx[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int b = 0; b < W; ++b) {
                        float Z0[3] = {0.0f, 0.0f, 0.0f}, Zd[3] = {0.0f, 0.0f, 0.0f}, Zx[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                        for (int c = 0; c < W; ++c) {
                            const int q = (c * T + b) * T + a;
                            const float wv[3] = {tq[q], tq[TN + q], tq[2 * TN + q]};
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                Z0[k] += Nz[c] * wv[k];
                                Zd[k] += dNz[c] * wv[k];
                                if (APIC) Zx[k] += zdx[c] * wv[k];
                            }
                        }
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            Y0[k] += Ny[b] * Z0[k];   // sum Ny Nz v
                            Yd[k] += dNy[b] * Z0[k];  // sum dNy Nz v
                            Yz[k] += Ny[b] * Zd[k];   // sum Ny dNz v
                            if (APIC) {
                                Yx[k] += ydx[b] * Z0[k];  // sum Ny dy Nz v
                                Yzx[k] += Ny[b] * Zx[k];  // sum Ny Nz dz v
                            }
                        }
                    }
                    const float xa = Nx[a], dxa = dNx[a];
                    vpic0 += xa * Y0[0]; vpic1 += xa * Y0[1]; vpic2 += xa * Y0[2];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        grad[k][0] += dxa * Y0[k];
                        grad[k][1] += xa * Yd[k];
                        grad[k][2] += xa * Yz[k];
                        if (APIC) {
                            Bm[k][0] += (xa * dxn[0][a]) * Y0[k];
                            Bm[k][1] += xa * Yx[k];
                            Bm[k][2] += xa * Yzx[k];
                        }
                    }
                }
                if (!APIC) {
                    // FLIP increment sum over v* - v, a second pass over the
                    // stencil so the main gather holds only the v* sums
#pragma unroll
                    for (int a = 0; a < W; ++a) {
                        float YF[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                        for (int b = 0; b < W; ++b) {
                            float ZF[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                            for (int c = 0; c < W; ++c) {
                                const int q = (c * T + b) * T + a;
#pragma unroll
                                for (int k = 0; k < 3; ++k) ZF[k] += Nz[c] * tq[(3 + k) * TN + q];
                            }
#pragma unroll
                            for (int k = 0; k < 3; ++k) YF[k] += Ny[b] * ZF[k];
                        }
                        const float xa = Nx[a];
                        vf0 += xa * YF[0]; vf1 += xa * YF[1]; vf2 += xa * YF[2];
                    }
                }
            }
            float n0, n1, n2;
            if (APIC) {
                n0 = vpic0; n1 = vpic1; n2 = vpic2;
            } else {
                const float ob = 1.0f - S.blend;
                n0 = S.blend * vf0 + ob * vpic0;
                n1 = S.blend * vf1 + ob * vpic1;
                n2 = S.blend * vf2 + ob * vpic2;
            }
            float Fe[3][3], Jp0, mu0, lam0, V0;
            if constexpr (PSTAGE) {
                // staged words have landed (waited for before this
                // particle's stores are issued, so the wait covers loads only)
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                const float* q = pst + (threadIdx.x & 63);
#pragma unroll
                for (int f = 0; f < 9; ++f) Fe[f / 3][f % 3] = q[f * 64];
                Jp0 = q[9 * 64];
                mu0 = q[10 * 64];
                lam0 = q[11 * 64];
                V0 = q[12 * 64];
            }
            // the material before the move list's atomic: a wait for it
            // would wait for the atomic too (memory returns count in order)
            const MaterialDev mat = P.mats[P.mat[p]];
            // Velocity, B, position and next sort key first: they depend on
            // the gather only, and storing them before the F update frees
            // their registers for the SVD.
            if (APIC) store_m3(P.B, N, p, Bm);
            if (BATCH ? S.particle_collision : pcollide)
                boundary_collision(S, x0 + S.dt * n0, x1 + S.dt * n1, x2 + S.dt * n2, n0, n1, n2);
            P.v[p] = n0;
            P.v[N + p] = n1;
            P.v[2 * N + p] = n2;
            if (advect) {
                const float y0 = x0 + S.dt * n0, y1 = x1 + S.dt * n1, y2 = x2 + S.dt * n2;
                P.x[p] = y0;
                P.x[N + p] = y1;
                P.x[2 * N + p] = y2;
                if (next_keys) {
                    knew = koff + cell_key(S, y0, y1, y2, Stencil<INTERP>::SHIFT, Stencil<INTERP>::W);
                    next_keys[pbase + p] = knew;
                    // uniform: the incremental sort is on; the entry is
                    // written after the F update
                    if (I.plan) ticket = move_begin(I, knew != kold);
                }
            }
            if constexpr (!PSTAGE) {
                load_m3(P.Fe, N, p, Fe);
                Jp0 = P.Jp[p];
                mu0 = P.mu0[p];
                lam0 = P.lam0[p];
                V0 = P.V0[p];
            }
            float Fn[3][3], E[6], J, tau[6];
            trial_deformation(Fe, grad, S.dt, Fn);
            const bool elastic = mat.model == MPM_MODEL_FIXED_COROTATED && elastic_trial(Fn, mat, E, J);
            if (__ballot(!elastic) == 0ull) {
                // every particle of the wave stays inside its thresholds:
                // Fe = Fn, Fp unchanged (neither read nor stored), stress
                // without an SVD from the cached det Fp
                elastic_fc_stress(E, J, Jp0, mu0, lam0, V0, mat, tau);
                store_m3(P.Fe, N, p, Fn);
            } else {
                // a plastic update somewhere in the wave: Fp from memory
                float Fp[3][3], U[3][3], sv[3], V[3][3];
                load_m3(P.Fp, N, p, Fp);
                update_deformation(Fn, Fe, Fp, mat, U, sv, V);
                // stress for the next P2G from the same decomposition
                const float Jp = m3_det(Fp);
                stress_from_svd(U, sv, V, Jp, mu0, lam0, V0, mat, tau, Fe);
                store_m3(P.Fe, N, p, Fe);
                store_m3(P.Fp, N, p, Fp);
                P.Jp[p] = Jp;
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) P.tau[(int64_t)k * N + p] = tau[k];
            if (I.plan && advect && next_keys) move_end(I, ticket, knew, pbase + p, i);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------- fused G2P+P2G
// APIC sums over the W^3 stencil of the v* tile: v = sum w v*, grad v =
// sum v* (grad w)^T, B = sum w v* (x_i - x_p)^T, as nested z/y/x sums (the
// gather of k_g2p's APIC path).
template <int W, int T, int TN>
__device__ __forceinline__ void apic_gather(const float* tq, const float (&Nx)[W], const float (&Ny)[W],
                                            const float (&Nz)[W], const float (&dNx)[W], const float (&dNy)[W],
                                            const float (&dNz)[W], const float (&dxn)[3][W], float (&v)[3],
                                            float (&grad)[3][3], float (&Bm)[3][3]) {
    float zdx[W], ydx[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        zdx[j] = Nz[j] * dxn[2][j];
        ydx[j] = Ny[j] * dxn[1][j];
    }
#pragma unroll
    for (int a = 0; a < W; ++a) {
        float Y0[3] = {0.0f, 0.0f, 0.0f}, Yd[3] = {0.0f, 0.0f, 0.0f}, Yz[3] = {0.0f, 0.0f, 0.0f};
        float Yx[3] = {0.0f, 0.0f, 0.0f}, Yzx[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < W; ++b) {
            float Z0[3] = {0.0f, 0.0f, 0.0f}, Zd[3] = {0.0f, 0.0f, 0.0f}, Zx[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const int q = (c * T + b) * T + a;
                const float wv[3] = {tq[q], tq[TN + q], tq[2 * TN + q]};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Z0[k] += Nz[c] * wv[k];
                    Zd[k] += dNz[c] * wv[k];
                    Zx[k] += zdx[c] * wv[k];
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Y0[k] += Ny[b] * Z0[k];
                Yd[k] += dNy[b] * Z0[k];
                Yz[k] += Ny[b] * Zd[k];
                Yx[k] += ydx[b] * Z0[k];
                Yzx[k] += Ny[b] * Zx[k];
            }
        }
        const float xa = Nx[a], dxa = dNx[a];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] += xa * Y0[k];
            grad[k][0] += dxa * Y0[k];
            grad[k][1] += xa * Yd[k];
            grad[k][2] += xa * Yz[k];
            Bm[k][0] += (xa * dxn[0][a]) * Y0[k];
            Bm[k][1] += xa * Yx[k];
            Bm[k][2] += xa * Yzx[k];
        }
    }
}

// One particle's folded-grid contributions to the nodes of its stencil
// (particlesToGrid, src/apic_mpm_solver3d.cu:245-304, with p + dt f summed
// per node): mass m w and m w (v + C (x_i - x_p)) + vs grad w, vs = dt tau.
// Tensor-product factors, (y, z) pairs outermost so that few of them are
// live at once.  With mip = m Nx[a] yz, cx(a) = v + C_0 dx(a) and
// cyz(b,c) = C_1 dy(b) + C_2 dz(c), the momentum plus force of a node,
//   mip (cx(a) + cyz(b,c)) + vs_0 dNx[a] yz + vs_1 Nx[a] dyz + vs_2 Nx[a] ydz,
// regroups by the node's x factor into
//   yz R(a) + Nx[a] Syz(b,c),
//   R(a)     = m Nx[a] cx(a) + vs_0 dNx[a]        (per particle and a)
//   Syz(b,c) = m yz cyz(b,c) + vs_1 dyz + vs_2 ydz  (per (b, c)),
// two operations per component and node; the mass is m Nx[a] yz.
// Each node's four values are summed over the lanes of the particle's cell
// (segmented scan) and the cell's last lane adds them to the tile at tq
// (the stencil's first node).  The read-add-write of the tile is pipelined
// by one node: a node's tile values are read before the next node's sums
// are computed, and added and written after, so the LDS latency overlaps
// arithmetic.  Writer lanes of different cells can reach one tile word
// through different nodes, so every lane's write of node k must precede any
// lane's read of node k + 1: the LDS accesses stay in node order (a wave's
// LDS instructions execute in order; the compiler barriers keep the
// compiler from reordering them).
template <int W, int T, int TN>
__device__ __forceinline__ void folded_cell_sums(float* tq, const SegPlan& sp, float m, const float (&vel)[3],
                                                 const float (&Cm)[3][3], const float (&vs)[3][3],
                                                 const float (&dxn)[3][W], const float (&Nx)[W],
                                                 const float (&Ny)[W], const float (&Nz)[W], const float (&dNx)[W],
                                                 const float (&dNy)[W], const float (&dNz)[W]) {
    float ma[W], R[W][3];
#pragma unroll
    for (int a = 0; a < W; ++a) {
        ma[a] = m * Nx[a];
#pragma unroll
        for (int k = 0; k < 3; ++k) R[a][k] = fmaf(ma[a], vel[k] + Cm[k][0] * dxn[0][a], vs[k][0] * dNx[a]);
    }
    float pv[4], po[4];  // the pending node: its sums and the tile values read for it
    int pq = 0;
#pragma unroll
    for (int b = 0; b < W; ++b)
#pragma unroll
        for (int c = 0; c < W; ++c) {
            const float yz = Ny[b] * Nz[c], dyz = dNy[b] * Nz[c], ydz = Ny[b] * dNz[c];
            const float myz = m * yz;
            float Syz[3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
                Syz[k] = fmaf(myz, Cm[k][1] * dxn[1][b] + Cm[k][2] * dxn[2][c], vs[k][1] * dyz + vs[k][2] * ydz);
#pragma unroll
            for (int a = 0; a < W; ++a) {
                float v[4];
                v[0] = ma[a] * yz;
#pragma unroll
                for (int k = 0; k < 3; ++k) v[1 + k] = fmaf(yz, R[a][k], Nx[a] * Syz[k]);
                seg_scan<4>(v, sp);
                const int q = (c * T + b) * T + a;
                const bool first = b == 0 && c == 0 && a == 0;
                asm volatile("" ::: "memory");
                if (sp.writer) {
                    if (!first) {
#pragma unroll
                        for (int f = 0; f < 4; ++f) tq[f * TN + pq] = po[f] + pv[f];
                    }
                    asm volatile("" ::: "memory");
#pragma unroll
                    for (int f = 0; f < 4; ++f) po[f] = tq[f * TN + q];
                }
#pragma unroll
                for (int f = 0; f < 4; ++f) pv[f] = v[f];
                pq = q;
            }
        }
    asm volatile("" ::: "memory");
    if (sp.writer) {
#pragma unroll
        for (int f = 0; f < 4; ++f) tq[f * TN + pq] = po[f] + pv[f];
    }
    asm volatile("" ::: "memory");
}

// Particles that changed cell (lanes set in `moved`, uniform over the wave):
// they are not in their new cell's segment, so the wave adds their folded
// contributions node by node.  It takes them one at a time: the particle's
// state is read from its lane (readlane: scalar registers) and each of the
// first W^3 lanes adds one stencil node, so a particle costs one pass
// instead of W^3 serial node passes.  A node inside the item's tile (ot, of
// T^3 nodes from (ox, oy, oz); null for no tile) is added into that LDS
// tile, which the item flushes with the rest; a particle that stayed in its
// key block has its whole stencil there.  Other nodes take global atomics.
template <int W>
__device__ __forceinline__ float pick(const float (&v)[W], int i) {
    float r = v[0];
#pragma unroll
    for (int j = 1; j < W; ++j) r = i == j ? v[j] : r;
    return r;
}
__device__ __forceinline__ float lane_value(float v, int src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}
template <int INTERP, int T>
__device__ __forceinline__ void scatter_moved(const SceneDev& S, float* grid, uint32_t* touched, float* ot, int ox,
                                              int oy, int oz, uint64_t moved, int lane, const float (&xn)[3], float m,
                                              const float (&vel)[3], const float (&Cm)[3][3],
                                              const float (&vs)[3][3]) {
    constexpr int W = Stencil<INTERP>::W;
    constexpr int TN = T * T * T;
    const int a = lane / (W * W), b = (lane / W) % W, c = lane % W;  // this lane's node (x, y, z offsets)
    while (moved) {
        const int src = __builtin_ctzll(moved);
        moved &= moved - 1;
        float x[3], v[3], C[3][3], t[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = lane_value(xn[k], src);
            v[k] = lane_value(vel[k], src);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                C[k][j] = lane_value(Cm[k][j], src);
                t[k][j] = lane_value(vs[k][j], src);
            }
        }
        const float pm = lane_value(m, src);
        if (lane >= W * W * W) continue;
        int bx, by, bz;
        float Nx[W], Ny[W], Nz[W], dNx[W], dNy[W], dNz[W];
        axis_weights<INTERP>(x[0], S.ox, S.h, S.inv_h, bx, Nx, dNx);
        axis_weights<INTERP>(x[1], S.oy, S.h, S.inv_h, by, Ny, dNy);
        axis_weights<INTERP>(x[2], S.oz, S.h, S.inv_h, bz, Nz, dNz);
        const int gi = bx + a, gj = by + b, gk = bz + c;
        if (!in_grid(S, gi, gj, gk)) continue;
        const float nx = pick<W>(Nx, a), ny = pick<W>(Ny, b), nz = pick<W>(Nz, c);
        const float w = (nx * ny) * nz;
        const float gw[3] = {pick<W>(dNx, a) * ny * nz, nx * pick<W>(dNy, b) * nz, nx * ny * pick<W>(dNz, c)};
        const float dx[3] = {(S.ox + S.h * (float)gi) - x[0], (S.oy + S.h * (float)gj) - x[1],
                             (S.oz + S.h * (float)gk) - x[2]};
        float val[4];
        val[0] = pm * w;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float ck = v[k] + ((C[k][0] * dx[0] + C[k][1] * dx[1]) + C[k][2] * dx[2]);
            const float fk = (t[k][0] * gw[0] + t[k][1] * gw[1]) + t[k][2] * gw[2];
            val[1 + k] = fmaf(val[0], ck, fk);
        }
        const int lx = gi - ox, ly = gj - oy, lz = gk - oz;
        if (ot && (unsigned)lx < (unsigned)T && (unsigned)ly < (unsigned)T && (unsigned)lz < (unsigned)T) {
            float* const q = ot + ((lz * T + ly) * T + lx);
#pragma unroll
            for (int f = 0; f < 4; ++f) atomicAdd(q + f * TN, val[f]);
        } else {
            const int64_t ni = node_index(S, gi, gj, gk);
#pragma unroll
            for (int f = 0; f < 4; ++f) unsafeAtomicAdd(&grid[(int64_t)f * S.plane + ni], val[f]);
            touched[storage_block(S, gi, gj, gk)] = 1u;
        }
    }
}

// gridToParticles of step n fused with particlesToGrid of step n+1 (APIC,
// folded grid, one scene): the item's v* tile is gathered as in k_g2p
// (grid update fused into the tile load, F update, stress, advection, next
// key); then, with the particle still in registers, its contributions to
// the next step's grid (ogrid/otouched, the other grid buffer) are summed
// per cell and added into the wave's own LDS tile, which the item flushes
// with one atomic per non-zero node value.  The reference runs the two
// transfers as separate passes with a fresh grid between them
// (src/apic_mpm_solver3d.cu:99-124); the sums are the same up to float
// order.  Lanes are consecutive sorted particles, so a cell's lanes are
// adjacent; a particle whose new key differs from its sorted key (a few per
// 10^4 on a slab, 1.5% on small fast scenes) keeps its lane in the cell's
// segment with zero contribution and the wave scatters it with global
// atomics, one node per lane.  The stress of the new state is used at
// once, so it is not cached.
#ifndef MPM_G2P2G_WAVES
#define MPM_G2P2G_WAVES 3
#endif
#ifndef MPM_G2P2G_WG
#define MPM_G2P2G_WG 128
#endif
// 3 waves per SIMD: the kernel needs ~162 VGPRs (at 128, 4 waves, ~24
// spill); its LDS (19.6 KB per 128-thread workgroup) would allow 4.
// BATCH: the items span the scenes of a batch; each item's scene, its
// second grid (ogrid, otouched) and its particle collision flag come from
// the table, and the next sort keys carry the scene's key offset.
template <int INTERP, int WG, bool BATCH>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(MPM_G2P2G_WAVES)))
k_g2p2g(SceneDev S0, ParticlesDev P0, ItemsDev I, const SceneEntry* __restrict__ tab, float* __restrict__ ogrid0,
        uint32_t* __restrict__ otouched0, int pcollide0, uint32_t* __restrict__ next_keys) {
    constexpr int NW = WG / 64;  // waves: each has its own output tile
    constexpr int W = Stencil<INTERP>::W;
    constexpr int T = 4 + W - 1;
    constexpr int TN = T * T * T;
    constexpr int TR = (TN + WG - 1) / WG;
    constexpr int PW = 13;
    __shared__ float tile[3 * TN];       // v* of this step
    __shared__ float stage[4 * TN];      // the next item's folded tile in flight
    __shared__ float pstage[PW * WG];    // Fe, Jp, mu0, lam0, V0 of each lane's particle
    __shared__ float otile[NW * 4 * TN];  // next step's sums, one tile per wave
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t nitems = *I.count;
    float* const ot = otile + wave * (4 * TN);
    ItemGeom gnext{};
    const XcdWalk walk = xcd_walk();
    const uint32_t it0 = xcd_item(walk, walk.k);
    if (it0 < nitems) {
        SceneDev Sn;
        ParticlesDev Pn;
        uint32_t pbn, kon;
        item_scene<BATCH>(S0, P0, I, tab, it0, Sn, Pn, pbn, kon);
        gnext = item_geom(Sn, I, it0, nitems, kon);
        fetch_folded_tile<T, TR, WG>(Sn, gnext, stage);
    }
    for (uint32_t k = walk.k, it = it0; it < nitems; k += walk.step, it = xcd_item(walk, k)) {
        SceneDev S;
        ParticlesDev P;
        uint32_t pbase, koff;
        item_scene<BATCH>(S0, P0, I, tab, it, S, P, pbase, koff);
        float MPM_GLOBAL* ogrid = (float MPM_GLOBAL*)ogrid0;
        uint32_t MPM_GLOBAL* otouched = (uint32_t MPM_GLOBAL*)otouched0;
        int pcollide = pcollide0;
        if constexpr (BATCH) {
            const uint32_t sc = I.scene[it];
            ogrid = tab[sc].grid2;
            otouched = tab[sc].touched2;
            pcollide = S.particle_collision;
        }
        const int64_t N = P.n;
        const ItemGeom g = gnext;
        if (!g.escape) commit_folded_tile<T, TR, WG>(S, g, stage, tile);
        for (int q = threadIdx.x; q < NW * 4 * TN; q += WG) otile[q] = 0.0f;
        __syncthreads();
        const uint32_t itn = xcd_item(walk, k + walk.step);
        if (itn < nitems) {
            SceneDev Sn;
            ParticlesDev Pn;
            uint32_t pbn, kon;
            item_scene<BATCH>(S0, P0, I, tab, itn, Sn, Pn, pbn, kon);
            gnext = item_geom(Sn, I, itn, nitems, kon);
            fetch_folded_tile<T, TR, WG>(Sn, gnext, stage);
        }
        // each wave takes every other 64-particle chunk of the item (no
        // barrier inside: a wave's trip count is its own, and every pass
        // has at least one particle); all lanes of a wave take part in the
        // cell sums
        for (uint32_t base = g.s + 64u * (uint32_t)wave; base < g.e; base += WG) {
            const uint32_t i = base + lane;
            const bool active = i < g.e;
            float xn[3] = {0.0f, 0.0f, 0.0f}, vn[3] = {0.0f, 0.0f, 0.0f}, Cn[3][3], vs[3][3], m = 0.0f;
            uint32_t kold = 0xFFFFFFFFu, knew = 0xFFFFFFFFu;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) { Cn[r][c] = 0.0f; vs[r][c] = 0.0f; }
            if (active) {
                // ---- step n: gridToParticles (src/apic_mpm_solver3d.cu:326-388)
                const uint32_t praw = I.perm[i], p = praw - pbase;
                kold = I.sorted_keys[i];
                if (I.snap_k) {  // the sorted order, kept for the next step's incremental sort
                    I.snap_k[i] = kold;
                    I.snap_p[i] = praw;
                }
                float* const pst = pstage + wave * (PW * 64);
#pragma unroll 1
                for (int f = 0; f < 9; ++f)
                    __builtin_amdgcn_global_load_lds((const void*)(P.Fe + (int64_t)f * N + p), (lds_void*)(pst + f * 64),
                                                     4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.Jp + p), (lds_void*)(pst + 9 * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.mu0 + p), (lds_void*)(pst + 10 * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.lam0 + p), (lds_void*)(pst + 11 * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((const void*)(P.V0 + p), (lds_void*)(pst + 12 * 64), 4, 0, 0);
                const float x0 = P.x[p], x1 = P.x[N + p], x2 = P.x[2 * N + p];
                m = P.m[p];
                float grad[3][3], Bm[3][3], v[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) { grad[r][c] = 0.0f; Bm[r][c] = 0.0f; }
                if (!g.escape) {
                    int bx, by, bz;
                    float Nx[W], Ny[W], Nz[W], dNx[W], dNy[W], dNz[W], dxn[3][W];
                    axis_weights<INTERP>(x0, S.ox, S.h, S.inv_h, bx, Nx, dNx);
                    axis_weights<INTERP>(x1, S.oy, S.h, S.inv_h, by, Ny, dNy);
                    axis_weights<INTERP>(x2, S.oz, S.h, S.inv_h, bz, Nz, dNz);
#pragma unroll
                    for (int j = 0; j < W; ++j) {
                        dxn[0][j] = (S.ox + S.h * (float)(bx + j)) - x0;
                        dxn[1][j] = (S.oy + S.h * (float)(by + j)) - x1;
                        dxn[2][j] = (S.oz + S.h * (float)(bz + j)) - x2;
                    }
                    const float* const tq = tile + (((bz - g.oz) * T + (by - g.oy)) * T + (bx - g.ox));
                    apic_gather<W, T, TN>(tq, Nx, Ny, Nz, dNx, dNy, dNz, dxn, v, grad, Bm);
                }
                store_m3(P.B, N, p, Bm);
                if (pcollide) boundary_collision(S, x0 + S.dt * v[0], x1 + S.dt * v[1], x2 + S.dt * v[2], v[0], v[1], v[2]);
                P.v[p] = v[0];
                P.v[N + p] = v[1];
                P.v[2 * N + p] = v[2];
                xn[0] = x0 + S.dt * v[0];
                xn[1] = x1 + S.dt * v[1];
                xn[2] = x2 + S.dt * v[2];
                P.x[p] = xn[0];
                P.x[N + p] = xn[1];
                P.x[2 * N + p] = xn[2];
                knew = koff + cell_key(S, xn[0], xn[1], xn[2], Stencil<INTERP>::SHIFT, Stencil<INTERP>::W);
                next_keys[pbase + p] = knew;
                // updateDeformationGradient and the stress of the new state
                // (the material loaded before the wait below, and the move
                // list's atomic issued after it: its result is used only
                // after the F update)
                const MaterialDev mat = P.mats[P.mat[p]];
                float Fe[3][3];
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // staged words have landed
                MoveTicket ticket{};
                if (I.plan) ticket = move_begin(I, knew != kold);
                float* const q = pst + lane;
#pragma unroll
                for (int f = 0; f < 9; ++f) Fe[f / 3][f % 3] = q[f * 64];
                const float Jp = q[9 * 64], mu0 = q[10 * 64], lam0 = q[11 * 64], V0 = q[12 * 64];
                // the lane's staging words are free now: B and the new
                // position wait there through the F update (fewer registers)
#pragma unroll
                for (int k = 0; k < 3; ++k) q[k * 64] = xn[k];
#pragma unroll
                for (int f = 0; f < 9; ++f) q[(3 + f) * 64] = Bm[f / 3][f % 3];
                float Fn[3][3], E[6], J, tau[6];
                trial_deformation(Fe, grad, S.dt, Fn);
                const bool elastic = mat.model == MPM_MODEL_FIXED_COROTATED && elastic_trial(Fn, mat, E, J);
                if (__ballot(!elastic) == 0ull) {
                    elastic_fc_stress(E, J, Jp, mu0, lam0, V0, mat, tau);
                    store_m3(P.Fe, N, p, Fn);
                } else {
                    // a plastic update somewhere in the wave: Fp from memory
                    float Fp[3][3], U[3][3], sv[3], V[3][3];
                    load_m3(P.Fp, N, p, Fp);
                    update_deformation(Fn, Fe, Fp, mat, U, sv, V);
                    const float Jn = m3_det(Fp);
                    stress_from_svd(U, sv, V, Jn, mu0, lam0, V0, mat, tau, Fe);
                    store_m3(P.Fe, N, p, Fe);
                    store_m3(P.Fp, N, p, Fp);
                    P.Jp[p] = Jn;
                }
                if (I.plan) move_end(I, ticket, knew, praw, i);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    vn[k] = v[k];
                    xn[k] = q[k * 64];
                }
#pragma unroll
                for (int f = 0; f < 9; ++f) Cn[f / 3][f % 3] = q[(3 + f) * 64] * S.dinv;
                const float t[6] = {S.dt * tau[0], S.dt * tau[1], S.dt * tau[2], S.dt * tau[3], S.dt * tau[4],
                                    S.dt * tau[5]};
                vs[0][0] = t[0]; vs[0][1] = t[1]; vs[0][2] = t[2];
                vs[1][0] = t[1]; vs[1][1] = t[3]; vs[1][2] = t[4];
                vs[2][0] = t[2]; vs[2][1] = t[4]; vs[2][2] = t[5];
            }
            // ---- step n+1: particlesToGrid into the next grid
            const bool moved = active && knew != kold;
            const uint64_t movers = __ballot(moved);
            if (movers)
                scatter_moved<INTERP, T>(S, ogrid, otouched, g.escape ? nullptr : ot, g.ox, g.oy, g.oz, movers, lane,
                                         xn, m, vn, Cn, vs);
            if (moved) {
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) vs[r][c] = 0.0f;  // no force term in the old cell's sums
            }
            if (!g.escape) {
                // segments by the sorted key: a moved lane stays in its old
                // cell's segment with zero weight
                const bool inscan = active && !moved;
                const uint32_t skey = active ? kold : 0xFFFFFFFFu;
                const uint32_t prev = __shfl_up(skey, 1, 64);
                const SegPlan sp = seg_plan(__ballot(lane == 0 || prev != skey), active, lane);
                int bx = 0, by = 0, bz = 0;
                float Nx[W], Ny[W], Nz[W], dNx[W], dNy[W], dNz[W], dxn[3][W];
                axis_weights<INTERP>(inscan ? xn[0] : 0.0f, S.ox, S.h, S.inv_h, bx, Nx, dNx);
                axis_weights<INTERP>(inscan ? xn[1] : 0.0f, S.oy, S.h, S.inv_h, by, Ny, dNy);
                axis_weights<INTERP>(inscan ? xn[2] : 0.0f, S.oz, S.h, S.inv_h, bz, Nz, dNz);
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    dxn[0][j] = (S.ox + S.h * (float)(bx + j)) - xn[0];
                    dxn[1][j] = (S.oy + S.h * (float)(by + j)) - xn[1];
                    dxn[2][j] = (S.oz + S.h * (float)(bz + j)) - xn[2];
                }
                // the segment's cell in the tile, from its key (cell bits)
                const int cxk = (int)(skey & 3u), cyk = (int)((skey >> 2) & 3u), czk = (int)((skey >> 4) & 3u);
                float* const tq = ot + ((czk * T + cyk) * T + cxk);
                folded_cell_sums<W, T, TN>(tq, sp, inscan ? m : 0.0f, vn, Cn, vs, dxn, Nx, Ny, Nz, dNx, dNy, dNz);
            }
        }
        __syncthreads();
        if (!g.escape) {
            for (int q = threadIdx.x; q < TN; q += WG) {
                const int lx = q % T, ly = (q / T) % T, lz = q / (T * T);
                const int gi = g.ox + lx, gj = g.oy + ly, gk = g.oz + lz;
                if (!in_grid(S, gi, gj, gk)) continue;
                const int64_t ni = node_index(S, gi, gj, gk);
                bool any = false;
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    float val = otile[f * TN + q];
#pragma unroll
                    for (int w = 1; w < NW; ++w) val += otile[(4 * w + f) * TN + q];
                    if (val != 0.0f) {
                        unsafeAtomicAdd(&ogrid[(int64_t)f * S.plane + ni], val);
                        any = true;
                    }
                }
                if (any) otouched[storage_block(S, gi, gj, gk)] = 1u;
            }
        }
        __syncthreads();
    }
}

void launch_g2p2g(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, int pcollide, float* ogrid,
                  uint32_t* otouched, uint32_t* next_keys, hipStream_t st) {
    if (I.n <= 0) return;
    MPM_PERSISTENT((k_g2p2g<MPM_INTERP_QUADRATIC_BSPLINE, MPM_G2P2G_WG, false>), MPM_G2P2G_WG, 0, st, S, P, I,
                   (const SceneEntry*)nullptr, ogrid, otouched, pcollide, next_keys);
}
void launch_g2p2g_batch(const SceneEntry* tab, const ItemsDev& I, int grid_wgs, uint32_t* next_keys,
                        hipStream_t st) {
    if (I.n <= 0) return;
    const SceneDev S0{};
    const ParticlesDev P0{};
    MPM_PERSISTENT((k_g2p2g<MPM_INTERP_QUADRATIC_BSPLINE, MPM_G2P2G_WG, true>), MPM_G2P2G_WG, grid_wgs, st, S0, P0,
                   I, tab, (float*)nullptr, (uint32_t*)nullptr, 0, next_keys);
}

template <int INTERP, int APIC, int SRC, bool BATCH>
static void g2p_dispatch_wg(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, const SceneEntry* tab,
                            int advect, int pcollide, int wgs, int wg, hipStream_t st, uint32_t* nk) {
    if (wg == 128)
        MPM_PERSISTENT((k_g2p<INTERP, APIC, SRC, BATCH, 128>), 128, wgs, st, S, P, I, tab, advect, pcollide, nk);
    else if (wg == 256)
        MPM_PERSISTENT((k_g2p<INTERP, APIC, SRC, BATCH, 256>), 256, wgs, st, S, P, I, tab, advect, pcollide, nk);
    else
        MPM_PERSISTENT((k_g2p<INTERP, APIC, SRC, BATCH, 64>), 64, wgs, st, S, P, I, tab, advect, pcollide, nk);
}
template <int INTERP, int APIC>
static void g2p_dispatch_src(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, const SceneEntry* tab,
                             int src, int advect, int pcollide, int wgs, int wg, hipStream_t st, uint32_t* nk) {
    if constexpr (APIC) {
        if (src == 3) {
            if (tab) g2p_dispatch_wg<INTERP, APIC, 3, true>(S, P, I, tab, advect, pcollide, wgs, wg, st, nk);
            else g2p_dispatch_wg<INTERP, APIC, 3, false>(S, P, I, tab, advect, pcollide, wgs, wg, st, nk);
            return;
        }
    }
    if (tab) g2p_dispatch_wg<INTERP, APIC, 0, true>(S, P, I, tab, advect, pcollide, wgs, wg, st, nk);
    else if (src == 0) g2p_dispatch_wg<INTERP, APIC, 0, false>(S, P, I, tab, advect, pcollide, wgs, wg, st, nk);
    else if (src == 1)
        MPM_PERSISTENT((k_g2p<INTERP, APIC, 1, false, kG2PWG>), kG2PWG, wgs, st, S, P, I, tab, advect, pcollide, nk);
    else
        MPM_PERSISTENT((k_g2p<INTERP, APIC, 2, false, kG2PWG>), kG2PWG, wgs, st, S, P, I, tab, advect, pcollide, nk);
}
template <int INTERP>
static void g2p_dispatch_apic(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, const SceneEntry* tab,
                              int apic, int src, int advect, int pcollide, int wgs, int wg, hipStream_t st,
                              uint32_t* nk) {
    if (apic) g2p_dispatch_src<INTERP, 1>(S, P, I, tab, src, advect, pcollide, wgs, wg, st, nk);
    else g2p_dispatch_src<INTERP, 0>(S, P, I, tab, src, advect, pcollide, wgs, wg, st, nk);
}
void launch_g2p(const SceneDev& S, const ParticlesDev& P, const ItemsDev& I, int interp, int apic, int src, int advect,
                int pcollide, int grid_wgs, hipStream_t st, uint32_t* next_keys, const SceneEntry* tab, int wg) {
    if (I.n <= 0) return;
    switch (interp) {
        case MPM_INTERP_LINEAR:
            g2p_dispatch_apic<MPM_INTERP_LINEAR>(S, P, I, tab, apic, src, advect, pcollide, grid_wgs, wg, st,
                                                 next_keys);
            break;
        case MPM_INTERP_QUADRATIC_BSPLINE:
            g2p_dispatch_apic<MPM_INTERP_QUADRATIC_BSPLINE>(S, P, I, tab, apic, src, advect, pcollide, grid_wgs, wg,
                                                            st, next_keys);
            break;
        default:
            g2p_dispatch_apic<MPM_INTERP_CUBIC_BSPLINE>(S, P, I, tab, apic, src, advect, pcollide, grid_wgs, wg, st,
                                                        next_keys);
            break;
    }
}

// ------------------------------------------------------------ stress cache
// Stress of the current particle state (after creation, initialize() or a
// state write); afterwards every G2P keeps it current.
__global__ void k_stress(ParticlesDev P) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int64_t N = P.n;
    float Fe[3][3], U[3][3], V[3][3], sv[3], tau[6];
    load_m3(P.Fe, N, (uint32_t)p, Fe);
    const MaterialDev mat = P.mats[P.mat[p]];
    svd3(Fe, U, sv, V);
    stress_from_svd(U, sv, V, P.Jp[p], P.mu0[p], P.lam0[p], P.V0[p], mat, tau, Fe);
    if (mat.model == MPM_MODEL_NEO_HOOKEAN) store_m3(P.Fe, N, (uint32_t)p, Fe);
#pragma unroll
    for (int k = 0; k < 6; ++k) P.tau[(int64_t)k * N + p] = tau[k];
}
void launch_stress(const ParticlesDev& P, hipStream_t st) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_stress, dim3((unsigned)((P.n + 127) / 128)), dim3(128), 0, st, P);
}

// ------------------------------------------------------------ grid passes
// Per-workgroup reduction of a wave-uniform count (lane 0 of each wave
// contributes), added with one atomic.
__device__ __forceinline__ void wg_count_add(unsigned long long* dst, unsigned long long mine) {
    __shared__ unsigned long long part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        if (t) atomicAdd(dst, t);
    }
}

// Grid clear: one lane per storage block reads the touched flags
// (coalesced); the wave then clears the touched blocks it found (64 nodes,
// one per lane) and counts their nodes with mass > 0 (the active nodes of
// the step that wrote them).  clear_block: one block (the batched clear).
__device__ __forceinline__ unsigned long long clear_block(float* grid, int64_t plane, uint32_t* touched, int64_t b,
                                                          int nfields, int lane) {
    const int64_t ni = b * 64 + lane;
    const unsigned long long c = (unsigned long long)__popcll(__ballot(grid[ni] > 0.0f));
    for (int f = 0; f < nfields; ++f) grid[(int64_t)f * plane + ni] = 0.0f;
    if (lane == 0) touched[b] = 0u;
    return c;
}

// Each wave takes the touched blocks among 64 consecutive storage blocks,
// kClearBatch at a time: their mass loads are all issued before the first
// count waits on one, so a wave pays one memory latency per batch rather
// than one per block.
constexpr int kClearBatch = 8;
__global__ void __launch_bounds__(256) k_clear(SceneDev S, int64_t nsb, int nfields, unsigned long long* active) {
    const int lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t todo = __ballot(b < nsb && S.touched[b] != 0u);
    const int64_t wave0 = b - lane;
    unsigned long long cnt = 0;
    while (todo) {
        int64_t blk[kClearBatch];
        float m[kClearBatch];
#pragma unroll
        for (int u = 0; u < kClearBatch; ++u) {
            blk[u] = -1;
            if (todo) {
                blk[u] = wave0 + __builtin_ctzll(todo);
                todo &= todo - 1;
            }
        }
#pragma unroll
        for (int u = 0; u < kClearBatch; ++u) m[u] = blk[u] >= 0 ? S.grid[blk[u] * 64 + lane] : 0.0f;
#pragma unroll
        for (int u = 0; u < kClearBatch; ++u) {
            if (blk[u] < 0) continue;
            cnt += (unsigned long long)__popcll(__ballot(m[u] > 0.0f));
            const int64_t ni = blk[u] * 64 + lane;
            for (int f = 0; f < nfields; ++f) S.grid[(int64_t)f * S.plane + ni] = 0.0f;
            if (lane == 0) S.touched[blk[u]] = 0u;
        }
    }
    if (active) wg_count_add(active, lane == 0 ? cnt : 0ull);
}

__global__ void __launch_bounds__(256) k_count_active(SceneDev S, int64_t nsb, unsigned long long* active) {
    const int lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t todo = __ballot(b < nsb && S.touched[b] != 0u);
    const int64_t wave0 = b - lane;
    unsigned long long cnt = 0;
    while (todo) {
        const int l = __builtin_ctzll(todo);
        todo &= todo - 1;
        cnt += (unsigned long long)__popcll(__ballot(S.grid[(wave0 + l) * 64 + lane] > 0.0f));
    }
    wg_count_add(active, lane == 0 ? cnt : 0ull);
}

// Batched clear over the storage blocks of every scene: global block index
// gb lies in scene s when block_start[s] <= gb < block_start[s + 1].  Each
// scene's active-node count goes to its own counter.
// mode 0: every scene's grid, its active nodes counted (the two-kernel
// step); mode 1: the grid of the scenes without pre-scattered sums, counted
// (the fused step's first step for them); mode 2: every scene's second
// grid, counted for the scenes whose last step was fused (it is the grid
// that step read) and not for the others (sums dropped when they left the
// fused state).
__device__ __forceinline__ float MPM_GLOBAL* clear_grid_of(const SceneEntry& e, int mode) {
    return mode == 2 ? e.grid2 : e.S.grid;
}
__device__ __forceinline__ uint32_t MPM_GLOBAL* clear_touched_of(const SceneEntry& e, int mode) {
    return mode == 2 ? e.touched2 : e.S.touched;
}
__global__ void __launch_bounds__(256) k_clear_batch(const SceneEntry* __restrict__ tab,
                                                     const int64_t* __restrict__ block_start, int nscenes,
                                                     int nfields, int mode) {
    const int lane = threadIdx.x & 63;
    const int64_t gb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = block_start[nscenes];
    int sc = 0;
    bool t = false;
    if (gb < total) {
        int lo = 0, hi = nscenes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (block_start[mid] <= gb) lo = mid;
            else hi = mid - 1;
        }
        sc = lo;
        const bool take = mode != 1 || tab[sc].prescat == 0u;
        t = take && clear_touched_of(tab[sc], mode)[gb - block_start[sc]] != 0u;
    }
    uint64_t todo = __ballot(t);
    // Blocks are in scene order, so a wave's touched blocks form a few runs
    // of one scene each: their counts are summed per run and added with one
    // atomic per run (one atomic per block on a single scene's counter
    // serialises at the L2).
    int cur = -1;
    unsigned long long run = 0;
    while (todo) {
        const int l = __builtin_ctzll(todo);
        todo &= todo - 1;
        const int s = __builtin_amdgcn_readfirstlane(__shfl(sc, l, 64));
        if (s != cur) {
            if (lane == 0 && run) atomicAdd(tab[cur].active, run);
            cur = s;
            run = 0;
        }
        const int64_t b = (gb - lane + l) - block_start[s];
        const unsigned long long c = clear_block(clear_grid_of(tab[s], mode), tab[s].S.plane,
                                                 clear_touched_of(tab[s], mode), b, nfields, lane);
        if (mode != 2 || tab[s].prescat != 0u) run += c;
    }
    if (lane == 0 && run) atomicAdd(tab[cur].active, run);
}

static unsigned grid_pass_wgs(int64_t nsb) {
    int64_t w = (nsb + 3) / 4;  // 4 waves per 256-thread workgroup
    if (w > 2048) w = 2048;
    if (w < 1) w = 1;
    return (unsigned)w;
}

void launch_clear(const SceneDev& S, int64_t nsb, int nfields, unsigned long long* active, hipStream_t st) {
    if (nsb <= 0) return;
    hipLaunchKernelGGL(k_clear, dim3((unsigned)((nsb + 255) / 256)), dim3(256), 0, st, S, nsb, nfields, active);
}
void launch_clear_batch(const SceneEntry* tab, const int64_t* block_start, int nscenes, int64_t total_blocks,
                        int nfields, hipStream_t st, int mode) {
    if (total_blocks <= 0 || nscenes <= 0) return;
    hipLaunchKernelGGL(k_clear_batch, dim3((unsigned)((total_blocks + 255) / 256)), dim3(256), 0, st, tab,
                       block_start, nscenes, nfields, mode);
}

// After a fused batch step: every scene's second grid (just scattered)
// becomes its grid, and the grid just read its second grid.
__global__ void k_swap_grids(SceneEntry* __restrict__ tab, int nscenes) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nscenes) return;
    float MPM_GLOBAL* g = tab[i].S.grid;
    uint32_t MPM_GLOBAL* t = tab[i].S.touched;
    tab[i].S.grid = tab[i].grid2;
    tab[i].S.touched = tab[i].touched2;
    tab[i].grid2 = g;
    tab[i].touched2 = t;
    tab[i].prescat = 1u;
}
void launch_swap_grids(SceneEntry* tab, int nscenes, hipStream_t st) {
    if (nscenes <= 0) return;
    hipLaunchKernelGGL(k_swap_grids, dim3((unsigned)((nscenes + 63) / 64)), dim3(64), 0, st, tab, nscenes);
}
void launch_count_active(const SceneDev& S, int64_t nsb, unsigned long long* active, hipStream_t st) {
    if (nsb <= 0) return;
    hipLaunchKernelGGL(k_count_active, dim3((unsigned)((nsb + 255) / 256)), dim3(256), 0, st, S, nsb, active);
}

// Unfused grid sub-steps over touched blocks (untouched nodes hold zero
// mass and zero v*, for which all three operations are no-ops).
__global__ void k_grid_op(SceneDev S, int64_t nsb, int op, float dt) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t b = wave; b < nsb; b += nwaves) {
        if (S.touched[b] == 0u) continue;
        const int64_t bx = b % S.nbx, by = (b / S.nbx) % S.nby, bz = b / ((int64_t)S.nbx * S.nby);
        const int i = (int)(4 * bx) + (lane & 3), j = (int)(4 * by) + ((lane >> 2) & 3), k = (int)(4 * bz) + (lane >> 4);
        if (i >= S.rx || j >= S.ry || k >= S.rz) continue;
        const int64_t ni = b * 64 + lane;
        float* G = S.grid;
        const int64_t pl = S.plane;
        const float m = G[ni];
        if (op == 0) {  // computeGravityForces, src/mpm_solver3d.cu:394-403
            if (m > MPM_FLT_EPS) G[(GF_FORCE + 1) * pl + ni] = fmaf(-S.gravity, m, G[(GF_FORCE + 1) * pl + ni]);
        } else if (op == 1) {  // updateVelocity, src/grid3d.cu:36-41
            if (m > MPM_FLT_EPS) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float v = G[(GF_VEL + a) * pl + ni] / m;
                    G[(GF_VEL + a) * pl + ni] = v;
                    G[(GF_VSTAR + a) * pl + ni] = v + (dt * G[(GF_FORCE + a) * pl + ni]) / m;
                }
            }
        } else {  // gridCollision, src/mpm_solver3d.cu:405-426
            float wx = G[(GF_VSTAR + 0) * pl + ni], wy = G[(GF_VSTAR + 1) * pl + ni], wz = G[(GF_VSTAR + 2) * pl + ni];
            const float gx = (S.ox + S.h * (float)i) + dt * wx;
            const float gy = (S.oy + S.h * (float)j) + dt * wy;
            const float gz = (S.oz + S.h * (float)k) + dt * wz;
            boundary_collision(S, gx, gy, gz, wx, wy, wz);
            G[(GF_VSTAR + 0) * pl + ni] = wx;
            G[(GF_VSTAR + 1) * pl + ni] = wy;
            G[(GF_VSTAR + 2) * pl + ni] = wz;
        }
    }
}
void launch_grid_op(const SceneDev& S, int64_t nsb, int op, float dt, hipStream_t st) {
    hipLaunchKernelGGL(k_grid_op, dim3(grid_pass_wgs(nsb)), dim3(256), 0, st, S, nsb, op, dt);
}

// ------------------------------------------------------- particle passes
__global__ void k_particles_collision(SceneDev S, ParticlesDev P, float dt) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int64_t N = P.n;
    float v0 = P.v[p], v1 = P.v[N + p], v2 = P.v[2 * N + p];
    boundary_collision(S, P.x[p] + dt * v0, P.x[N + p] + dt * v1, P.x[2 * N + p] + dt * v2, v0, v1, v2);
    P.v[p] = v0; P.v[N + p] = v1; P.v[2 * N + p] = v2;
}
void launch_particles_collision(const SceneDev& S, const ParticlesDev& P, float dt, hipStream_t st) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_particles_collision, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, st, S, P, dt);
}
__global__ void k_advect(ParticlesDev P, float dt) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int64_t N = P.n;
#pragma unroll
    for (int a = 0; a < 3; ++a) P.x[a * N + p] = P.x[a * N + p] + dt * P.v[a * N + p];
}
void launch_advect(const ParticlesDev& P, float dt, hipStream_t st) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_advect, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, st, P, dt);
}

// ----------------------------------------------------- layout conversion
__global__ void k_pack(HostLayoutDev H, ParticlesDev P) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const int64_t N = P.n;
    const int64_t id = P.pid[s];
    P.V0[s] = H.volume0[id];
    P.m[s] = H.mass[id];
    P.mu0[s] = H.mu0[id];
    P.lam0[s] = H.lam0[id];
    P.mat[s] = H.mat[id];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        P.x[a * N + s] = H.pos[3 * id + a];
        P.v[a * N + s] = H.vel[3 * id + a];
    }
    float Fp[3][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        P.Fe[k * N + s] = H.Fe[9 * id + k];
        Fp[k / 3][k % 3] = H.Fp[9 * id + k];
        P.Fp[k * N + s] = Fp[k / 3][k % 3];
        if (P.B) P.B[k * N + s] = H.B ? H.B[9 * id + k] : 0.0f;
    }
    P.Jp[s] = m3_det(Fp);
}
void launch_pack(const HostLayoutDev& H, const ParticlesDev& P, hipStream_t st) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, st, H, P);
}
__global__ void k_unpack(ParticlesDev P, HostLayoutDev H) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const int64_t N = P.n;
    const int64_t id = P.pid[s];
    H.volume0[id] = P.V0[s];
    H.mass[id] = P.m[s];
    H.mu0[id] = P.mu0[s];
    H.lam0[id] = P.lam0[s];
    H.mat[id] = P.mat[s];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        H.pos[3 * id + a] = P.x[a * N + s];
        H.vel[3 * id + a] = P.v[a * N + s];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        H.Fe[9 * id + k] = P.Fe[k * N + s];
        H.Fp[9 * id + k] = P.Fp[k * N + s];
        if (H.B) H.B[9 * id + k] = P.B ? P.B[k * N + s] : 0.0f;
    }
}
void launch_unpack(const ParticlesDev& P, const HostLayoutDev& H, hipStream_t st) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, st, P, H);
}
__global__ void k_positions4(ParticlesDev P, float* out) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.n) return;
    const int64_t N = P.n;
    const int64_t id = P.pid[s];
    out[4 * id + 0] = P.x[s];
    out[4 * id + 1] = P.x[N + s];
    out[4 * id + 2] = P.x[2 * N + s];
    out[4 * id + 3] = 1.0f;
}
void launch_positions4(const ParticlesDev& P, float* out, hipStream_t st) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_positions4, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, st, P, out);
}
// Physical reorder: dst slot i <- src slot perm[i].
__global__ void k_permute(ParticlesDev S, ParticlesDev D, const uint32_t* __restrict__ perm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const int64_t N = S.n;
    const int64_t s = perm[i];
    D.V0[i] = S.V0[s];
    D.m[i] = S.m[s];
    D.mu0[i] = S.mu0[s];
    D.lam0[i] = S.lam0[s];
    D.Jp[i] = S.Jp[s];
    D.mat[i] = S.mat[s];
    D.pid[i] = S.pid[s];
#pragma unroll
    for (int k = 0; k < 6; ++k) D.tau[k * N + i] = S.tau[k * N + s];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        D.x[a * N + i] = S.x[a * N + s];
        D.v[a * N + i] = S.v[a * N + s];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        D.Fe[k * N + i] = S.Fe[k * N + s];
        D.Fp[k * N + i] = S.Fp[k * N + s];
        if (S.B) D.B[k * N + i] = S.B[k * N + s];
    }
}
void launch_permute(const ParticlesDev& src, const ParticlesDev& dst, const uint32_t* perm, hipStream_t st) {
    if (src.n <= 0) return;
    hipLaunchKernelGGL(k_permute, dim3((unsigned)((src.n + 255) / 256)), dim3(256), 0, st, src, dst, perm);
}
// ------------------------------------------------------ store reorder
// The particle store keeps the cell order of its last reorder while the
// particles move; the slots a wave reaches through the sort permutation
// spread out (8.4M slab: G2P 0.88 -> 1.19 ms after 1000 steps, back to 0.87
// once the store is re-sorted).  A reorder permutes the store in place into
// the current sorted order, one field at a time through a scratch buffer
// (slot i <- slot perm[i]); for a batch every member's store is permuted
// into its range of the global sorted order.

// Field f of a store (planar components): 0 x, 1 v, 2 B, 3 Fe, 4 Fp,
// 5 tau, 6 m, 7 V0, 8 mu0, 9 lambda0, 10 Jp, 11 pid (moved as 32-bit
// words); kStoreFields - 1 = mat is 16-bit.  Null for B without APIC.
constexpr int kStoreFields = 13;
__device__ __forceinline__ uint32_t MPM_GLOBAL* store_field32(const ParticlesDev& P, int f, int& w) {
    float MPM_GLOBAL* p = nullptr;
    w = 1;
    switch (f) {
        case 0: w = 3; p = P.x; break;
        case 1: w = 3; p = P.v; break;
        case 2: w = 9; p = P.B; break;
        case 3: w = 9; p = P.Fe; break;
        case 4: w = 9; p = P.Fp; break;
        case 5: w = 6; p = P.tau; break;
        case 6: p = P.m; break;
        case 7: p = P.V0; break;
        case 8: p = P.mu0; break;
        case 9: p = P.lam0; break;
        case 10: p = P.Jp; break;
        default: return P.pid;
    }
    return reinterpret_cast<uint32_t MPM_GLOBAL*>(p);
}

// Member of a batch that owns global slot g (members hold consecutive
// slot ranges in table order).
__device__ __forceinline__ int slot_scene(const SceneEntry* __restrict__ tab, int nscenes, int64_t g) {
    int lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].pstart <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Scene of every global slot of a batch (one binary search per slot, once
// per reorder instead of once per field and pass).
__global__ void __launch_bounds__(256) k_slot_scenes(const SceneEntry* __restrict__ tab, int nscenes, int64_t ntot,
                                                     uint32_t* __restrict__ slot_sc) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ntot) slot_sc[g] = (uint32_t)slot_scene(tab, nscenes, g);
}

template <bool BATCH, bool BACK>
__global__ void __launch_bounds__(256) k_reorder_field(ParticlesDev P0, const SceneEntry* __restrict__ tab,
                                                       const uint32_t* __restrict__ slot_sc,
                                                       const uint32_t* __restrict__ perm, int64_t ntot, int f,
                                                       uint32_t* __restrict__ scratch) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    // gather: scratch[i] <- slot perm[i]; back: slot i <- scratch[i]
    const int64_t g = BACK ? i : (int64_t)perm[i];
    ParticlesDev P = P0;
    int64_t ps = 0;
    if (BATCH) {
        // slot g = perm[i] lies in the same scene as position i: scenes hold
        // ascending, disjoint key ranges in slot order, so the sorted order
        // keeps each scene's particles in its own slot range; reading the
        // scene at i keeps that load coalesced and off the perm[i] chain
        const uint32_t sc = slot_sc[i];
        P = tab[sc].P;
        ps = tab[sc].pstart;
    }
    const int64_t l = g - ps, n = P.n;
    if (f == kStoreFields - 1) {
        uint16_t* s16 = reinterpret_cast<uint16_t*>(scratch);
        if (BACK) P.mat[l] = s16[i]; else s16[i] = P.mat[l];
        return;
    }
    int w;
    uint32_t MPM_GLOBAL* base = store_field32(P, f, w);
    if (!base) return;
    for (int c = 0; c < w; ++c) {
        if (BACK) base[c * n + l] = scratch[c * ntot + i];
        else scratch[c * ntot + i] = base[c * n + l];
    }
}

void launch_reorder_store(const ParticlesDev& P, const SceneEntry* tab, int nscenes, const uint32_t* perm,
                          int64_t ntot, uint32_t* scratch, bool apic, hipStream_t st, uint32_t* slot_sc) {
    if (ntot <= 0) return;
    const dim3 grid((unsigned)((ntot + 255) / 256)), block(256);
    if (tab) hipLaunchKernelGGL(k_slot_scenes, grid, block, 0, st, tab, nscenes, ntot, slot_sc);
    for (int f = 0; f < kStoreFields; ++f) {
        if (f == 2 && !apic) continue;
        if (tab) {
            hipLaunchKernelGGL((k_reorder_field<true, false>), grid, block, 0, st, P, tab, slot_sc, perm, ntot, f,
                               scratch);
            hipLaunchKernelGGL((k_reorder_field<true, true>), grid, block, 0, st, P, tab, slot_sc, perm, ntot, f,
                               scratch);
        } else {
            hipLaunchKernelGGL((k_reorder_field<false, false>), grid, block, 0, st, P, tab, nullptr, perm, ntot, f,
                               scratch);
            hipLaunchKernelGGL((k_reorder_field<false, true>), grid, block, 0, st, P, tab, nullptr, perm, ntot, f,
                               scratch);
        }
    }
}

// ---------------------------------------------------- batch reorder by chunk
// A batch spreads its store reorder over steps: its members are grouped
// into chunks of consecutive members (about ntot / chunks particles each,
// positions [bounds[k], bounds[k + 1]) of the global order), and each step
// permutes the members of chunk *rr into the step's sorted order, then
// advances *rr.  The index lives on the device, so a captured step does the
// same on every replay.  A chunk's positions are exactly its members' slots
// (a scene's keys are one contiguous range), so the gather pass copies every
// word of slot perm[pos] into a scratch row (kStoreWords rows of stride
// words, and the 16-bit material index beside them) and the scatter pass
// writes row i back to slot pos and sets perm[pos] = pos: the rest of the
// step then reads the reordered store through the identity.
constexpr int kStoreWords = 45;  // x 3, v 3, B 9, Fe 9, Fp 9, Jp, tau 6, m, V0, mu0, lam0, pid
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_reorder_chunk(const SceneEntry* __restrict__ tab,
                                                       const uint32_t* __restrict__ slot_sc, uint32_t* perm,
                                                       const int64_t* __restrict__ bounds, const uint32_t* rr,
                                                       int64_t stride, uint32_t* __restrict__ scratch, int apic) {
    const uint32_t k = *rr;
    const int64_t lo = bounds[k], hi = bounds[k + 1];
    uint16_t* const s16 = reinterpret_cast<uint16_t*>(scratch + (int64_t)kStoreWords * stride);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lo + i < hi;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = lo + i;
        const uint32_t sc = slot_sc[pos];
        const int64_t ps = tab[sc].pstart;
        const ParticlesDev& P = tab[sc].P;
        const int64_t n = P.n;
        const int64_t l = SCATTER ? pos - ps : (int64_t)perm[pos] - ps;
        int w = 0;
        auto move = [&](float MPM_GLOBAL* f, int comps) {
            for (int c = 0; c < comps; ++c, ++w) {
                uint32_t MPM_GLOBAL* const a = reinterpret_cast<uint32_t MPM_GLOBAL*>(f) + (int64_t)c * n + l;
                if (SCATTER) *a = scratch[(int64_t)w * stride + i];
                else scratch[(int64_t)w * stride + i] = *a;
            }
        };
        move(P.x, 3);
        move(P.v, 3);
        if (apic) move(P.B, 9);
        else w += 9;
        move(P.Fe, 9);
        move(P.Fp, 9);
        move(P.Jp, 1);
        move(P.tau, 6);
        move(P.m, 1);
        move(P.V0, 1);
        move(P.mu0, 1);
        move(P.lam0, 1);
        move(reinterpret_cast<float MPM_GLOBAL*>(P.pid), 1);
        if (SCATTER) {
            P.mat[l] = s16[i];
            perm[pos] = (uint32_t)pos;
        } else {
            s16[i] = P.mat[l];
        }
    }
}
__global__ void k_chunk_advance(uint32_t* rr, uint32_t nchunks) {
    if (threadIdx.x == 0) *rr = *rr + 1u < nchunks ? *rr + 1u : 0u;
}
void launch_reorder_chunk(const SceneEntry* tab, const uint32_t* slot_sc, uint32_t* perm, const int64_t* bounds,
                          uint32_t* rr, int nchunks, int64_t chunk_max, uint32_t* scratch, bool apic, hipStream_t st) {
    if (nchunks <= 0 || chunk_max <= 0) return;
    const int64_t want = (chunk_max + 255) / 256;
    const unsigned blocks = (unsigned)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(k_reorder_chunk<false>, dim3(blocks), dim3(256), 0, st, tab, slot_sc, perm, bounds, rr,
                       chunk_max, scratch, apic ? 1 : 0);
    hipLaunchKernelGGL(k_reorder_chunk<true>, dim3(blocks), dim3(256), 0, st, tab, slot_sc, perm, bounds, rr,
                       chunk_max, scratch, apic ? 1 : 0);
    hipLaunchKernelGGL(k_chunk_advance, dim3(1), dim3(64), 0, st, rr, (uint32_t)nchunks);
}
void launch_slot_scenes(const SceneEntry* tab, int nscenes, int64_t ntot, uint32_t* slot_sc, hipStream_t st) {
    if (ntot <= 0) return;
    hipLaunchKernelGGL(k_slot_scenes, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st, tab, nscenes, ntot,
                       slot_sc);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t* dst,
                             int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
void launch_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t* dst, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, idx, dst, n);
}

// Dense reference-order grid copy (linear index x + rx*(y + ry*z)).
__global__ void k_grid_dense(SceneDev S, float* mass, float* vel, float* vstar, float* force, int have_vstar) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)S.rx * S.ry * S.rz;
    if (i >= total) return;
    const int x = (int)(i % S.rx), y = (int)((i / S.rx) % S.ry), z = (int)(i / ((int64_t)S.rx * S.ry));
    const int64_t ni = node_index(S, x, y, z);
    const float* G = S.grid;
    const int64_t pl = S.plane;
    if (mass) mass[i] = G[ni];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (vel) vel[3 * i + a] = G[(GF_VEL + a) * pl + ni];
        if (force) force[3 * i + a] = G[(GF_FORCE + a) * pl + ni];
        if (vstar) vstar[3 * i + a] = have_vstar ? G[(GF_VSTAR + a) * pl + ni] : 0.0f;
    }
}
void launch_grid_dense(const SceneDev& S, float* mass, float* vel, float* vstar, float* force, int have_vstar,
                       hipStream_t st) {
    const int64_t total = (int64_t)S.rx * S.ry * S.rz;
    hipLaunchKernelGGL(k_grid_dense, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, S, mass, vel, vstar,
                       force, have_vstar);
}

// ------------------------------------------------------- density raster
// writeToOpenVDB's density (src/mpm_solver3d.cu:510-563) is the mass-only
// P2G of the current positions times 1/h^3; these kernels read it out of a
// one-plane block-major grid D (D.grid = sum of m w per node).
__global__ void k_density_dense(SceneDev D, float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)D.rx * D.ry * D.rz;
    if (i >= total) return;
    const int x = (int)(i % D.rx), y = (int)((i / D.rx) % D.ry), z = (int)(i / ((int64_t)D.rx * D.ry));
    out[i] = D.grid[node_index(D, x, y, z)] * D.inv_cell_vol;
}
void launch_density_dense(const SceneDev& D, float* out, hipStream_t st) {
    const int64_t total = (int64_t)D.rx * D.ry * D.rz;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_density_dense, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, D, out);
}

__global__ void k_touched_flags(const uint32_t* __restrict__ touched, int64_t n, uint32_t* __restrict__ flags) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = touched[i] != 0u ? 1u : 0u;
}
void launch_touched_flags(const uint32_t* touched, int64_t n, uint32_t* flags, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_touched_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, touched, n, flags);
}

// One 64-lane wave per storage block: a touched block b of rank r (exclusive
// scan of the touched flags) writes its node origin to origin[3r..3r+2] and
// its 64 densities (x fastest) to values[64r..64r+63]; ranks >= max_blocks
// are dropped (the caller sizes the output from the count).
__global__ void __launch_bounds__(256) k_density_blocks(SceneDev D, const uint32_t* __restrict__ touched,
                                                        const uint32_t* __restrict__ rank, int64_t nsb,
                                                        int32_t* __restrict__ origin, float* __restrict__ values,
                                                        int64_t max_blocks) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = t >> 6;
    const int l = (int)(t & 63);
    if (b >= nsb || touched[b] == 0u) return;
    const int64_t r = rank[b];
    if (r >= max_blocks) return;
    if (values) values[r * 64 + l] = D.grid[b * 64 + l] * D.inv_cell_vol;
    if (origin && l < 3) {
        const int64_t bx = b % D.nbx, by = (b / D.nbx) % D.nby, bz = b / ((int64_t)D.nbx * D.nby);
        origin[3 * r + l] = 4 * (int32_t)(l == 0 ? bx : (l == 1 ? by : bz));
    }
}
void launch_density_blocks(const SceneDev& D, const uint32_t* touched, const uint32_t* rank, int64_t nsb,
                           int32_t* origin, float* values, int64_t max_blocks, hipStream_t st) {
    if (nsb <= 0 || max_blocks <= 0) return;
    hipLaunchKernelGGL(k_density_blocks, dim3((unsigned)((nsb * 64 + 255) / 256)), dim3(256), 0, st, D, touched,
                       rank, nsb, origin, values, max_blocks);
}

// Largest |v|^2 over the particles, as the bits of a non-negative float
// (ordered like unsigned integers; a NaN sorts above every number).
__global__ void __launch_bounds__(256) k_max_speed2(ParticlesDev P, unsigned int* __restrict__ out) {
    float m = 0.0f;
    const int64_t N = P.n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const float a = P.v[i], b = P.v[N + i], c = P.v[2 * N + i];
        const float q = (a * a + b * b) + c * c;
        m = (q > m || q != q) ? q : m;
    }
    unsigned int u = __float_as_uint(m);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned int o = __shfl_xor(u, off, 64);
        u = o > u ? o : u;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, u);
}
void launch_max_speed2(const ParticlesDev& P, unsigned int* out, hipStream_t st) {
    if (P.n <= 0) return;
    int64_t w = (P.n + 255) / 256;
    if (w > 2048) w = 2048;
    hipLaunchKernelGGL(k_max_speed2, dim3((unsigned)w), dim3(256), 0, st, P, out);
}

__global__ void __launch_bounds__(256) k_count_escaped(const uint32_t* __restrict__ sk, int64_t n, uint32_t esc,
                                                       unsigned long long* out) {
    unsigned long long cnt = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < n;
         i += (int64_t)gridDim.x * blockDim.x)
        cnt += (unsigned long long)__popcll(__ballot(i < n && sk[i] == esc));
    wg_count_add(out, (threadIdx.x & 63) == 0 ? cnt : 0ull);
}
void launch_count_escaped(const uint32_t* sorted_keys, int64_t n, uint32_t escape_key, unsigned long long* out,
                          hipStream_t st) {
    if (n <= 0) return;
    int64_t w = (n + 255) / 256;
    if (w > 1024) w = 1024;
    hipLaunchKernelGGL(k_count_escaped, dim3((unsigned)w), dim3(256), 0, st, sorted_keys, n, escape_key, out);
}

// Health of every scene of a batch after its steps, one pass over the
// global slots: per scene the largest |v|^2 (float bits; a NaN wins), the
// particles with a non-finite position or velocity component, and the
// particles whose next key is the scene's escape key (stencil outside the
// grid).  keys: the global sort input the last G2P wrote (key offset +
// cell key per slot), so a slot's scene is the last whose key offset is <=
// its key.  out: 3 words per scene, zeroed by the caller.
__global__ void __launch_bounds__(256) k_batch_health(const SceneEntry* __restrict__ tab,
                                                      const uint32_t* __restrict__ key_offsets, int ns, int64_t ntot,
                                                      const uint32_t* __restrict__ keys, uint32_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < ntot; i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        const bool act = i < ntot;
        int sc = 0;
        uint32_t q = 0, bad = 0, esc = 0;
        if (act) {
            const uint32_t key = keys[i];
            int lo = 0, hi = ns - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (key_offsets[mid] <= key) lo = mid;
                else hi = mid - 1;
            }
            sc = lo;
            const int64_t p = i - tab[sc].pstart;
            const int64_t N = tab[sc].P.n;
            const float* x = tab[sc].P.x;
            const float* v = tab[sc].P.v;
            const float a = v[p], b = v[N + p], c = v[2 * N + p];
            const float x0 = x[p], x1 = x[N + p], x2 = x[2 * N + p];
            const float sq = (a * a + b * b) + c * c;
            q = __float_as_uint(sq) & 0x7FFFFFFFu;  // |v|^2 >= 0; a NaN keeps its payload above +inf
            bad = (isfinite(a) && isfinite(b) && isfinite(c) && isfinite(x0) && isfinite(x1) && isfinite(x2)) ? 0u : 1u;
            esc = key - key_offsets[sc] == tab[sc].S.escape_key ? 1u : 0u;
        }
        // a wave of one scene reduces first; scene boundaries add per lane
        const int s0 = __shfl(sc, 0, 64);
        if (__ballot(act && sc != s0) == 0ull) {
            uint32_t m = q, nb = bad, ne = esc;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint32_t o = __shfl_xor(m, off, 64);
                m = o > m ? o : m;
                nb += __shfl_xor(nb, off, 64);
                ne += __shfl_xor(ne, off, 64);
            }
            if (lane == 0) {
                atomicMax(&out[3 * s0], m);
                if (nb) atomicAdd(&out[3 * s0 + 1], nb);
                if (ne) atomicAdd(&out[3 * s0 + 2], ne);
            }
        } else if (act) {
            atomicMax(&out[3 * sc], q);
            if (bad) atomicAdd(&out[3 * sc + 1], bad);
            if (esc) atomicAdd(&out[3 * sc + 2], esc);
        }
    }
}
void launch_batch_health(const SceneEntry* tab, const uint32_t* key_offsets, int ns, int64_t ntot,
                         const uint32_t* keys, uint32_t* out, hipStream_t st) {
    if (ntot <= 0 || ns <= 0) return;
    int64_t w = (ntot + 255) / 256;
    if (w > 2048) w = 2048;
    hipLaunchKernelGGL(k_batch_health, dim3((unsigned)w), dim3(256), 0, st, tab, key_offsets, ns, ntot, keys, out);
}

__global__ void k_debug_svd(const float* A, int64_t n, float* U, float* s, float* V) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a[3][3], u[3][3], v[3][3], sv[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[r][c] = A[9 * i + 3 * r + c];
    svd3(a, u, sv, v);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        s[3 * i + r] = sv[r];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            U[9 * i + 3 * r + c] = u[r][c];
            V[9 * i + 3 * r + c] = v[r][c];
        }
    }
}
void launch_debug_svd(const float* A, int64_t n, float* U, float* s, float* V, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_debug_svd, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, A, n, U, s, V);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}

}  // namespace mpm
