// incsort.hip -- the per-step cell sort, updated instead of redone.
//
// Between two steps few particles change cell (0.05% of the 8.4M-particle
// slab per step, about 1% of a batch of small scenes).  The sorted
// order of step n+1 -- the stable sort of the new keys, i.e. ascending
// (key, slot) -- is the sorted order of step n with the moved particles
// taken out and put back at their new places:
//   * the transfer kernel copies step n's sorted keys and slots (the
//     snapshot) as it reads them;
//   * G2P appends each particle whose new key differs from its old one to a
//     list (new key, slot, old sorted position; kMoveLists sub-lists);
//   * here every list entry finds its place among the snapshot (lb: the
//     number of snapshot entries before it in (key, slot) order) by a binary
//     search, and the removed positions and insertion places are bucketed
//     by snapshot tile (counts, one scan, then an append per entry);
//   * one pass over the snapshot tiles finishes the job: each workgroup
//     ranks its tile's few removals and insertions in LDS and writes every
//     unmoved entry at i - (removed before i) + (inserted at or before i)
//     and every inserted entry at lb - (removed before lb) + (inserted
//     before it).
// The result is bit-identical to a stable radix sort of the new keys.  The
// device-side plan words pick this path or the full radix sort each step
// (the other path's launches return at once): the full sort runs when the
// list overflowed, a tile drew more entries than its merge stages, or the
// snapshot is not the previous step's (first step, state written, a
// single scene's store reordered, a batch's layout changed).  The step's item count resets the words afterwards
// (launch_items with plan).  Launch order within a step: launch_inc_prepare,
// the radix sort (which reads the decision), launch_inc_merge.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blockscan.h"
#include "launch.h"

namespace mpm {

constexpr int kIncWG = 256;

// (key, slot) order of two sort entries.
__device__ __forceinline__ bool entry_less(uint32_t ka, uint32_t sa, uint32_t kb, uint32_t sb) {
    return ka < kb || (ka == kb && sa < sb);
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

__global__ void k_inc_set(uint32_t* plan, uint32_t snap_ok) {
    if (threadIdx.x < kMoveLists) plan[kPlanLists + threadIdx.x * kMoveCountStride] = 0u;
    if (threadIdx.x != 0) return;
    plan[kPlanSnapOk] = snap_ok;
    plan[kPlanOverflow] = 0u;
}

// The entries of the sub-lists: blockIdx.x % kMoveLists picks the sub-list,
// the other blocks of the grid with that remainder stride over its filled
// part (a sub-list's count is at most its capacity unless it overflowed,
// and then no update runs).  j: the entry's index in the list arrays.
#define FOR_LIST_ENTRIES(plan, j)                                                                             \
    const uint32_t r_ = blockIdx.x % (uint32_t)kMoveLists;                                                    \
    const uint32_t c_ = (plan)[kPlanLists + r_ * kMoveCountStride];                                           \
    const uint32_t base_ = r_ * (plan)[kPlanCap];                                                             \
    const uint32_t step_ = (gridDim.x / (uint32_t)kMoveLists) * kIncWG;                                       \
    for (uint32_t q_ = (blockIdx.x / (uint32_t)kMoveLists) * kIncWG + threadIdx.x, j = base_ + q_; q_ < c_; \
         q_ += step_, j = base_ + q_)

// Each list entry's place among the snapshot, by binary search (a chain
// of dependent loads, the entries all in flight at once; in a batch within
// the entry's member), and the tile counts of its old position and of its
// place (zero on entry: the merge clears them after use).
__global__ void __launch_bounds__(kIncWG) k_inc_locate(const uint32_t* __restrict__ plan, MoveList L,
                                                       const uint32_t* __restrict__ snap_k,
                                                       const uint32_t* __restrict__ snap_p, int64_t n,
                                                       uint32_t tiles) {
    if (!plan_incremental(plan)) return;
    FOR_LIST_ENTRIES(plan, j) {
        const uint32_t k = L.key[j], s = L.slot[j];
        uint32_t lo = 0, hi = (uint32_t)n;  // the place lies in [lo, hi]
        if (L.nranges) {  // a batch: within its member's range (a short search in a cached table first)
            uint32_t a = 0, b = L.nranges;
            while (b - a > 1) {
                const uint32_t mid = (a + b) >> 1;
                if (L.range_start[mid] <= s) a = mid;
                else b = mid;
            }
            lo = L.range_start[a];
            hi = L.range_start[a + 1];
        }
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (entry_less(snap_k[mid], snap_p[mid], k, s)) lo = mid + 1;
            else hi = mid;
        }
        L.lb[j] = lo;
        atomicAdd(&L.tcount[L.pos[j] / (uint32_t)kIncTile], 1u);
        atomicAdd(&L.tcount[tiles + 1u + lo / (uint32_t)kIncTile], 1u);
    }
}

// Bucket offsets: exclusive scans of the two count arrays (tiles + 1
// entries each) into the bucket cursors, by one workgroup per array (each
// thread a run of consecutive counts); a tile with more entries than the
// merge stages sends this step to the full sort.
__global__ void __launch_bounds__(kScanWG) k_inc_offsets(uint32_t* __restrict__ plan, MoveList L, uint32_t tiles) {
    if (!plan_incremental(plan)) return;
    __shared__ uint32_t ws[kScanWG / 64];
    const uint32_t m = tiles + 1u;
    const uint32_t* const cnt = L.tcount + (size_t)blockIdx.x * m;
    uint32_t* const cur = L.tnext + (size_t)blockIdx.x * m;
    const uint32_t per = (m + kScanWG - 1) / kScanWG;
    const uint32_t t0 = threadIdx.x * per < m ? threadIdx.x * per : m;
    const uint32_t t1 = t0 + per < m ? t0 + per : m;
    uint32_t sum = 0;
    bool over = false;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t v = cnt[t];
        over = over || v > (uint32_t)kIncStage;
        sum += v;
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(sum, ws, tot);
    for (uint32_t t = t0; t < t1; ++t) {
        cur[t] = run;
        run += cnt[t];
    }
    if (over) plan[kPlanOverflow] = 1u;
}

// Each entry's old position and entry index appended to the buckets of
// their tiles (the order inside a bucket is arbitrary: the merge ranks it).
__global__ void __launch_bounds__(kIncWG) k_inc_bucket(const uint32_t* __restrict__ plan, MoveList L,
                                                       uint32_t tiles) {
    if (!plan_incremental(plan)) return;
    FOR_LIST_ENTRIES(plan, j) {
        const uint32_t pos = L.pos[j], lb = L.lb[j];
        L.del_b[atomicAdd(&L.tnext[pos / (uint32_t)kIncTile], 1u)] = pos;
        L.ins_b[atomicAdd(&L.tnext[tiles + 1u + lb / (uint32_t)kIncTile], 1u)] = j;
    }
}

// One snapshot tile [i0, i1) per workgroup (the last one, past n when n is
// a multiple of the tile, takes the insertions at n); it also zeroes the
// tile's counts, also on a step the update does not run (after an overflow
// the counts are left over).  After the bucketing
// the cursors sit at the end of each bucket, so a bucket starts at
// cursor - count.  Removed positions are distinct and list entries have
// distinct (key, slot), so ranking by counting gives each its own place.
__global__ void __launch_bounds__(kIncWG) k_inc_merge(const uint32_t* __restrict__ plan, MoveList L,
                                                      const uint32_t* __restrict__ snap_k,
                                                      const uint32_t* __restrict__ snap_p, int64_t n, uint32_t tiles,
                                                      uint32_t* __restrict__ out_k, uint32_t* __restrict__ out_p) {
    const uint32_t t = blockIdx.x;
    if (!plan_incremental(plan)) {  // an overflow after the counting: clear them for the next step
        if (threadIdx.x == 0) {
            L.tcount[t] = 0u;
            L.tcount[tiles + 1u + t] = 0u;
        }
        return;
    }
    __shared__ uint32_t dr[kIncStage], sdel[kIncStage];              // removed positions: as bucketed, ascending
    __shared__ uint32_t ir[kIncStage], ik[kIncStage], is[kIncStage];  // insertions: place, key, slot
    __shared__ uint32_t slb[kIncStage];                                // insertion places, ascending
    const int64_t i0 = (int64_t)t * kIncTile;
    const int64_t i1 = i0 + kIncTile < n ? i0 + kIncTile : n;
    const uint32_t dc = L.tcount[t], rc = L.tcount[tiles + 1u + t];  // <= kIncStage (else overflow)
    const uint32_t d0 = L.tnext[t] - dc, r0 = L.tnext[tiles + 1u + t] - rc;
    for (uint32_t q = threadIdx.x; q < dc; q += kIncWG) dr[q] = L.del_b[d0 + q];
    for (uint32_t q = threadIdx.x; q < rc; q += kIncWG) {
        const uint32_t j = L.ins_b[r0 + q];
        ir[q] = L.lb[j];
        ik[q] = L.key[j];
        is[q] = L.slot[j];
    }
    __syncthreads();  // every thread has read the counts
    if (threadIdx.x == 0) {  // zero for the next step's counting
        L.tcount[t] = 0u;
        L.tcount[tiles + 1u + t] = 0u;
    }
    for (uint32_t q = threadIdx.x; q < dc; q += kIncWG) {
        const uint32_t v = dr[q];
        uint32_t r = 0;
        for (uint32_t x = 0; x < dc; ++x) r += dr[x] < v ? 1u : 0u;
        sdel[r] = v;
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < rc; q += kIncWG) {
        const uint32_t k = ik[q], s = is[q], lb = ir[q];
        uint32_t r = 0;  // (key, slot) order is also the order of the places
        for (uint32_t x = 0; x < rc; ++x) r += entry_less(ik[x], is[x], k, s) ? 1u : 0u;
        slb[r] = lb;
        const uint32_t dst = lb - (d0 + count_below(sdel, dc, lb)) + r0 + r;
        if ((int64_t)dst < n) {  // always, for a list that matches the snapshot
            out_k[dst] = k;
            out_p[dst] = s;
        }
    }
    __syncthreads();
    // consecutive lanes take consecutive entries: loads and (shifted) stores stay contiguous
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kIncWG) {
        const uint32_t iu = (uint32_t)i;
        const uint32_t dl = count_below(sdel, dc, iu);
        if (dl < dc && sdel[dl] == iu) continue;  // moved
        const uint32_t dst = iu - (d0 + dl) + r0 + count_below(slb, rc, iu + 1u);  // places <= i
        if ((int64_t)dst < n) {
            out_k[dst] = snap_k[i];
            out_p[dst] = snap_p[i];
        }
    }
}

uint32_t inc_tiles(int64_t n) { return (uint32_t)(n / kIncTile + 1); }

void launch_inc_set(uint32_t* plan, uint32_t snap_ok, hipStream_t st) {
    hipLaunchKernelGGL(k_inc_set, dim3(1), dim3(64), 0, st, plan, snap_ok);
}

void launch_inc_prepare(uint32_t* plan, const MoveList& L, const uint32_t* snap_k, const uint32_t* snap_p, int64_t n,
                        hipStream_t st) {
    const uint32_t tiles = inc_tiles(n);
    // 16 workgroups per sub-list
    hipLaunchKernelGGL(k_inc_locate, dim3(16 * kMoveLists), dim3(kIncWG), 0, st, plan, L, snap_k, snap_p, n, tiles);
    hipLaunchKernelGGL(k_inc_offsets, dim3(2), dim3(kScanWG), 0, st, plan, L, tiles);
}

void launch_inc_merge(const uint32_t* plan, const MoveList& L, const uint32_t* snap_k, const uint32_t* snap_p,
                      int64_t n, uint32_t* out_k, uint32_t* out_p, hipStream_t st) {
    const uint32_t tiles = inc_tiles(n);
    hipLaunchKernelGGL(k_inc_bucket, dim3(16 * kMoveLists), dim3(kIncWG), 0, st, plan, L, tiles);
    hipLaunchKernelGGL(k_inc_merge, dim3(tiles), dim3(kIncWG), 0, st, plan, L, snap_k, snap_p, n, tiles, out_k,
                       out_p);
}

}  // namespace mpm
