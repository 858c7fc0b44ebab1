// incsort.hip -- the per-step cell sort of a large scene, updated instead of
// redone.
//
// Between two steps few particles change cell (0.05% of the 8.4M-particle
// slab per step).  The sorted order of step n+1 -- the stable sort of the
// new keys, i.e. ascending (key, slot) -- is the sorted order of step n
// with the moved particles taken out and put back at their new places:
//   * the transfer kernel copies step n's sorted keys and permutation (the
//     snapshot) as it reads them;
//   * G2P appends each particle whose new key differs from its old one to a
//     short list (new key, slot, old sorted position);
//   * here the list is ranked by (key, slot) (by counting: it is short), each
//     entry finds its place among the snapshot by a 64-ary search, and one
//     pass over the snapshot moves every unmoved entry by
//     (#list entries placed before it) - (#moved entries before it).
// The result is bit-identical to a stable radix sort of the new keys.  The
// device-side plan words pick this path or the full radix sort each step
// (the other path's launches return at once): the full sort runs when the
// list overflowed (more than kMoveMax moves) or the snapshot is not the
// previous step's (first step, state written, store reordered).  The step's
// item count resets the words afterwards (launch_items with plan).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"

namespace mpm {

constexpr int kIncWG = 256;
constexpr int kIncTile = 2048;   // snapshot entries per merge workgroup
constexpr int kIncStage = 1024;  // list entries staged per merge workgroup

// (key, slot) order of two sort entries.
__device__ __forceinline__ bool entry_less(uint32_t ka, uint32_t sa, uint32_t kb, uint32_t sb) {
    return ka < kb || (ka == kb && sa < sb);
}

__global__ void k_inc_set(uint32_t* plan, uint32_t snap_ok) {
    if (threadIdx.x != 0) return;
    plan[kPlanMoved] = 0u;
    plan[kPlanSnapOk] = snap_ok;
}

// Each list entry's place among the snapshot (the number of snapshot
// entries before it in (key, slot) order), one wave per entry: every round
// the 64 lanes probe evenly spaced entries of the current range and the
// ballot of "probe < (key, slot)" (a prefix of the lanes: the snapshot is
// sorted) narrows the range ~65-fold -- 4 rounds of dependent loads for 8.4M
// entries instead of 23 for a binary search.
__global__ void __launch_bounds__(kIncWG) k_inc_locate(const uint32_t* __restrict__ plan, MoveList L,
                                                       const uint32_t* __restrict__ snap_k,
                                                       const uint32_t* __restrict__ snap_p, int64_t n) {
    if (!plan_incremental(plan)) return;
    const uint32_t c = plan[kPlanMoved];
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (kIncWG / 64);
    for (uint32_t j = blockIdx.x * (kIncWG / 64) + (threadIdx.x >> 6); j < c; j += waves) {  // wave-uniform
        const uint32_t k = L.key[j], s = L.slot[j];
        int64_t lo = 0, hi = n;  // the place lies in [lo, hi]
        for (int round = 0; round < 64 && lo < hi; ++round) {
            const int64_t span = hi - lo;
            const int64_t p = lo + span * (lane + 1) / 65;  // in [lo, hi)
            const int below = __popcll(__ballot(entry_less(snap_k[p], snap_p[p], k, s)));
            // lanes 0 .. below-1 probed entries before the place, lane `below` one at or after it
            const int64_t nlo = below > 0 ? lo + span * below / 65 + 1 : lo;
            const int64_t nhi = below < 64 ? lo + span * (below + 1) / 65 : hi;
            lo = nlo;
            hi = nhi;
        }
        if (lane == 0) L.lb[j] = (uint32_t)lo;
    }
}

// Counting over the list: rank = entries before this one in (key, slot)
// order; drank = entries whose old position is smaller; dbelow = old
// positions below this entry's place.  blockIdx.y takes one slice of the
// other entries and stores its partial counts (summed by k_inc_place).
__global__ void __launch_bounds__(kIncWG) k_inc_count(const uint32_t* __restrict__ plan, MoveList L) {
    if (!plan_incremental(plan)) return;
    const uint32_t c = plan[kPlanMoved];
    const uint32_t j0 = blockIdx.x * kIncWG;
    if (j0 >= c) return;
    const uint32_t per = (c + kIncSlices - 1) / kIncSlices;
    const uint32_t e0 = blockIdx.y * per, e1 = e0 + per < c ? e0 + per : c;
    __shared__ uint32_t sk[kIncWG], ss[kIncWG], sp[kIncWG];
    const uint32_t j = j0 + threadIdx.x;
    const bool own = j < c;
    const uint32_t k = own ? L.key[j] : 0u, s = own ? L.slot[j] : 0u;
    const uint32_t pos = own ? L.pos[j] : 0u, lb = own ? L.lb[j] : 0u;
    uint32_t rank = 0, drank = 0, dbelow = 0;
    for (uint32_t b = e0; b < e1; b += kIncWG) {
        const uint32_t e = b + threadIdx.x;
        if (e < e1) {
            sk[threadIdx.x] = L.key[e];
            ss[threadIdx.x] = L.slot[e];
            sp[threadIdx.x] = L.pos[e];
        }
        __syncthreads();
        const uint32_t m = e1 - b < (uint32_t)kIncWG ? e1 - b : (uint32_t)kIncWG;
#pragma unroll 8
        for (uint32_t q = 0; q < m; ++q) {
            rank += entry_less(sk[q], ss[q], k, s) ? 1u : 0u;
            drank += sp[q] < pos ? 1u : 0u;
            dbelow += sp[q] < lb ? 1u : 0u;
        }
        __syncthreads();
    }
    if (own) {
        uint32_t* const part = L.part + (size_t)blockIdx.y * kMoveMax + j;
        part[0] = rank;
        part[(size_t)kIncSlices * kMoveMax] = drank;
        part[(size_t)2 * kIncSlices * kMoveMax] = dbelow;
    }
}

// Moved entries to their final place; the sorted lists of removed
// positions and of insertion places for the merge.
__global__ void __launch_bounds__(kIncWG) k_inc_place(const uint32_t* __restrict__ plan, MoveList L, int64_t n,
                                                      uint32_t* __restrict__ out_k, uint32_t* __restrict__ out_p) {
    if (!plan_incremental(plan)) return;
    const uint32_t c = plan[kPlanMoved];
    const uint32_t j = blockIdx.x * kIncWG + threadIdx.x;
    if (j >= c) return;
    uint32_t rank = 0, drank = 0, dbelow = 0;
#pragma unroll
    for (int y = 0; y < kIncSlices; ++y) {
        const uint32_t* const part = L.part + (size_t)y * kMoveMax + j;
        rank += part[0];
        drank += part[(size_t)kIncSlices * kMoveMax];
        dbelow += part[(size_t)2 * kIncSlices * kMoveMax];
    }
    const uint32_t lb = L.lb[j];
    const uint32_t dst = rank + lb - dbelow;  // < n for a list that matches the snapshot
    if ((int64_t)dst < n) {
        out_k[dst] = L.key[j];
        out_p[dst] = L.slot[j];
    }
    if (drank < c) L.del_sorted[drank] = L.pos[j];
    if (rank < c) L.lb_sorted[rank] = lb;
}

// Number of entries of a[0..n) (ascending) that are < v.
__device__ __forceinline__ uint32_t count_below(const uint32_t* a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One pass over the snapshot: each entry that did not move goes to
// i - (moved entries before i) + (list entries placed at or before i).
// The few removed positions and insertion places of the tile are staged
// in LDS (global binary search when a tile holds more than kIncStage).
// Consecutive lanes take consecutive entries, so that both the loads and
// the (shifted) stores of a wave stay contiguous.
__global__ void __launch_bounds__(kIncWG) k_inc_merge(const uint32_t* __restrict__ plan, MoveList L,
                                                      const uint32_t* __restrict__ snap_k,
                                                      const uint32_t* __restrict__ snap_p, int64_t n,
                                                      uint32_t* __restrict__ out_k, uint32_t* __restrict__ out_p) {
    if (!plan_incremental(plan)) return;
    const uint32_t c = plan[kPlanMoved];
    __shared__ uint32_t bounds[4];
    __shared__ uint32_t sdel[kIncStage], slb[kIncStage];
    const int64_t i0 = (int64_t)blockIdx.x * kIncTile;
    const int64_t i1 = i0 + kIncTile < n ? i0 + kIncTile : n;
    if (threadIdx.x < 4) {
        const uint32_t* a = threadIdx.x < 2 ? L.del_sorted : L.lb_sorted;
        const uint32_t v = (uint32_t)((threadIdx.x & 1) ? i1 : i0);
        bounds[threadIdx.x] = count_below(a, c, v);
    }
    __syncthreads();
    const uint32_t d0 = bounds[0], d1 = bounds[1], r0 = bounds[2], r1 = bounds[3];
    const bool staged = d1 - d0 <= (uint32_t)kIncStage && r1 - r0 <= (uint32_t)kIncStage;
    if (staged) {
        for (uint32_t q = threadIdx.x; q < d1 - d0; q += kIncWG) sdel[q] = L.del_sorted[d0 + q];
        for (uint32_t q = threadIdx.x; q < r1 - r0; q += kIncWG) slb[q] = L.lb_sorted[r0 + q];
    }
    __syncthreads();
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kIncWG) {
        const uint32_t iu = (uint32_t)i;
        uint32_t d, r;
        bool removed;
        if (staged) {
            const uint32_t dl = count_below(sdel, d1 - d0, iu);
            removed = dl < d1 - d0 && sdel[dl] == iu;
            d = d0 + dl;
            r = r0 + count_below(slb, r1 - r0, iu + 1u);  // places <= i
        } else {
            d = count_below(L.del_sorted, c, iu);
            removed = d < c && L.del_sorted[d] == iu;
            r = count_below(L.lb_sorted, c, iu + 1u);
        }
        if (removed) continue;
        const uint32_t dst = iu - d + r;  // < n for a list that matches the snapshot
        if ((int64_t)dst < n) {
            out_k[dst] = snap_k[i];
            out_p[dst] = snap_p[i];
        }
    }
}

void launch_inc_set(uint32_t* plan, uint32_t snap_ok, hipStream_t st) {
    hipLaunchKernelGGL(k_inc_set, dim3(1), dim3(64), 0, st, plan, snap_ok);
}

void launch_inc_sort(const uint32_t* plan, const MoveList& L, const uint32_t* snap_k, const uint32_t* snap_p,
                     int64_t n, uint32_t* out_k, uint32_t* out_p, hipStream_t st) {
    const unsigned entry_blocks = (unsigned)((kMoveMax + kIncWG - 1) / kIncWG);
    // 4096 waves: one entry each up to 4096 moves (the slab moves ~4000
    // particles a step), a few each beyond
    hipLaunchKernelGGL(k_inc_locate, dim3(1024), dim3(kIncWG), 0, st, plan, L, snap_k, snap_p, n);
    hipLaunchKernelGGL(k_inc_count, dim3(entry_blocks, kIncSlices), dim3(kIncWG), 0, st, plan, L);
    hipLaunchKernelGGL(k_inc_place, dim3(entry_blocks), dim3(kIncWG), 0, st, plan, L, n, out_k, out_p);
    const unsigned tiles = (unsigned)((n + kIncTile - 1) / kIncTile);
    hipLaunchKernelGGL(k_inc_merge, dim3(tiles), dim3(kIncWG), 0, st, plan, L, snap_k, snap_p, n, out_k, out_p);
}

}  // namespace mpm
