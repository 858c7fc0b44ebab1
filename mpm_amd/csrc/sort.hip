// sort.hip -- stable LSD radix sort of (key, value) uint32 pairs and a
// device-wide exclusive scan, written for wave64 (gfx950).
//
// One pass over a digit (6 to 9 bits, planned by radix_passes) is three
// launches:
//   hist    : per-tile digit histogram, stored digit-major (digit * tiles + tile)
//   scan    : exclusive scan of that table -> global base of (digit, tile)
//   scatter : per-tile stable ranking (64-key chunks ranked with 8 ballots,
//             chunk prefix sums in LDS), staged in LDS in digit order, then
//             written out so that runs of one digit are contiguous stores.
// Stability follows from ranking by (tile, chunk, lane) = input order.
// No look-back spinning: every wave runs to completion independently.
// Every kernel loops over its tiles (workgroup b takes tiles b, b + grid,
// ...), so a sort can be launched on fewer workgroups than tiles: a sort
// whose launches mostly return at once (the plan words chose the
// incremental update) then costs fewer workgroup dispatches.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blockscan.h"
#include "launch.h"

namespace mpm {

constexpr int kMaxRadixBits = 9;
constexpr int kMaxRadix = 1 << kMaxRadixBits;
constexpr int kSortWG = 256;
constexpr int kSortIPT = 8;
constexpr int kSortTile = kSortWG * kSortIPT;  // 2048 keys per tile
constexpr int kSortChunks = kSortTile / 64;    // 32 chunks of 64 keys
constexpr int kChunksPerWave = kSortChunks / (kSortWG / 64);
constexpr int kLightGrid = 512;  // workgroups of a sort launched with few_passes

template <int BITS>
__device__ __forceinline__ uint64_t digit_match(uint32_t digit, uint64_t active) {
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        bool bit = (digit >> b) & 1u;
        uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

template <int BITS>
__global__ void __launch_bounds__(kSortWG) k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                        uint32_t mask, uint32_t* __restrict__ hist, int64_t tiles,
                                                        const uint32_t* __restrict__ plan) {
    if (plan && plan_incremental(plan)) return;
    constexpr int kRadix = 1 << BITS;
    __shared__ uint32_t h[kRadix];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        for (int i = tid; i < kRadix; i += kSortWG) h[i] = 0;
        __syncthreads();
        const int64_t base = tile * kSortTile;
        // all of the wave's key loads are issued before the first is used
        uint32_t kk[kChunksPerWave];
#pragma unroll
        for (int c = 0; c < kChunksPerWave; ++c) {
            const int64_t idx = base + (int64_t)(wave * kChunksPerWave + c) * 64 + lane;
            kk[c] = idx < n ? keys[idx] : 0u;
        }
#pragma unroll
        for (int c = 0; c < kChunksPerWave; ++c) {
            const int64_t idx = base + (int64_t)(wave * kChunksPerWave + c) * 64 + lane;
            const bool valid = idx < n;
            const uint64_t active = __ballot(valid);
            if (active == 0) continue;  // wave-uniform: the tail of the last tile
            const uint32_t digit = (kk[c] >> shift) & mask;
            const uint64_t m = digit_match<BITS>(digit, active);
            const int leader = __ffsll((unsigned long long)m) - 1;
            if (valid && lane == leader) atomicAdd(&h[digit], (uint32_t)__popcll(m));
        }
        __syncthreads();
        for (int d = tid; d < kRadix; d += kSortWG) hist[(int64_t)d * tiles + tile] = h[d];
        __syncthreads();
    }
}

template <int BITS, bool RAW>
__global__ void __launch_bounds__(kSortWG) k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                           const uint32_t* __restrict__ vals_in,
                                                           uint32_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out, int64_t n, int shift,
                                                           uint32_t mask, const uint32_t* __restrict__ hist_scan,
                                                           int64_t tiles, const uint32_t* __restrict__ plan) {
    if (plan && plan_incremental(plan)) return;
    constexpr int kRadix = 1 << BITS;
    // digits per thread: 2 (9 bits), 1 (8 bits), or 1 for the first kRadix
    // threads and none for the rest (narrower digits)
    constexpr int DPT = kRadix >= kSortWG ? kRadix / kSortWG : 1;
    static_assert(kRadix < kSortWG || DPT * kSortWG == kRadix, "radix must be a multiple of the workgroup size");
    const bool owns = kRadix >= kSortWG || (int)threadIdx.x < kRadix;
    __shared__ uint16_t chunk_off[kSortChunks][kRadix];  // 16 KB (8 bits) or 32 KB (9 bits)
    __shared__ uint32_t digit_base[kRadix];
    __shared__ uint32_t global_base[kRadix];
    __shared__ uint32_t stage_k[kSortTile];
    __shared__ uint32_t stage_v[kSortTile];
    __shared__ uint32_t wave_sums[kSortWG / 64];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t base = tile * kSortTile;
    const int tile_count = (int)((n - base) < kSortTile ? (n - base) : kSortTile);

    uint32_t* co32 = reinterpret_cast<uint32_t*>(&chunk_off[0][0]);
    for (int i = tid; i < kSortChunks * kRadix / 2; i += kSortWG) co32[i] = 0u;  // kRadix >= 2
    __syncthreads();

    uint32_t my_key[kChunksPerWave], my_val[kChunksPerWave], my_digit[kChunksPerWave], my_rank[kChunksPerWave];
#pragma unroll
    for (int c = 0; c < kChunksPerWave; ++c) {
        int chunk = wave * kChunksPerWave + c;
        int64_t idx = base + (int64_t)chunk * 64 + lane;
        bool valid = idx < n;
        uint32_t k = valid ? keys_in[idx] : 0u;
        uint32_t v = valid ? vals_in[idx] : 0u;
        uint32_t digit = (k >> shift) & mask;
        uint64_t active = __ballot(valid);
        uint64_t m = digit_match<BITS>(digit, active);
        uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        my_rank[c] = (uint32_t)__popcll(m & lt);
        int leader = __ffsll((unsigned long long)m) - 1;
        if (valid && lane == leader) chunk_off[chunk][digit] = (uint16_t)__popcll(m);
        my_key[c] = k;
        my_val[c] = v;
        my_digit[c] = valid ? digit : 0xFFFFFFFFu;
    }
    __syncthreads();

    // Per digit: exclusive prefix over chunks (thread t owns digits
    // t*DPT .. t*DPT+DPT-1, so digit order follows thread order).
    uint32_t dtot[DPT];
    uint32_t total = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        dtot[j] = 0;
        if (!owns) continue;
        const int d = tid * DPT + j;
        uint32_t acc = 0;
        for (int c = 0; c < kSortChunks; ++c) {
            uint32_t cnt = chunk_off[c][d];
            chunk_off[c][d] = (uint16_t)acc;
            acc += cnt;
        }
        dtot[j] = acc;
        total += acc;
        if (!RAW) global_base[d] = hist_scan[(int64_t)d * tiles + tile];
    }
    if (RAW) {
        // few tiles: the digit table is not scanned beforehand; each
        // workgroup forms its bases from the raw counts, as the sum over
        // smaller digits of their totals plus this digit's earlier tiles
        uint32_t gtot[DPT];
        uint32_t gsum = 0;
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            gtot[j] = 0;
            if (!owns) continue;
            const int d = tid * DPT + j;
            uint32_t before = 0, all = 0;
            for (int64_t t = 0; t < tiles; ++t) {
                const uint32_t c = hist_scan[(int64_t)d * tiles + t];
                before += t < tile ? c : 0u;
                all += c;
            }
            global_base[d] = before;
            gtot[j] = all;
            gsum += all;
        }
        uint32_t gall;
        uint32_t run = block_exclusive_scan(gsum, wave_sums, gall);
        if (owns) {
#pragma unroll
            for (int j = 0; j < DPT; ++j) {
                global_base[tid * DPT + j] += run;
                run += gtot[j];
            }
        }
    }
    // Block exclusive scan of the per-thread digit totals.
    uint32_t incl = total;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t up = __shfl_up(incl, off, 64);
        if (lane >= off) incl += up;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    uint32_t wave_prefix = 0;
    for (int w = 0; w < wave; ++w) wave_prefix += wave_sums[w];
    if (owns) {
        uint32_t run = wave_prefix + incl - total;
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            digit_base[tid * DPT + j] = run;
            run += dtot[j];
        }
    }
    __syncthreads();

#pragma unroll
    for (int c = 0; c < kChunksPerWave; ++c) {
        if (my_digit[c] == 0xFFFFFFFFu) continue;
        int chunk = wave * kChunksPerWave + c;
        uint32_t d = my_digit[c];
        uint32_t pos = digit_base[d] + chunk_off[chunk][d] + my_rank[c];
        stage_k[pos] = my_key[c];
        stage_v[pos] = my_val[c];
    }
    __syncthreads();

    for (int i = tid; i < tile_count; i += kSortWG) {
        uint32_t k = stage_k[i];
        uint32_t d = (k >> shift) & mask;
        int64_t out = (int64_t)global_base[d] + (i - (int64_t)digit_base[d]);
        keys_out[out] = k;
        vals_out[out] = stage_v[i];
    }
    __syncthreads();  // the next tile reuses every LDS array
    }
}

// ------------------------------------------------------------------ scan
// (kScanWG, kScanIPT, block_exclusive_scan: blockscan.h)

__global__ void __launch_bounds__(kScanWG) k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                         uint32_t* __restrict__ partial,
                                                         const uint32_t* __restrict__ plan) {
    if (plan && plan_incremental(plan)) return;
    __shared__ uint32_t ws[kScanWG / 64];
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        int64_t base = tile * kScanTile;
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < kScanIPT; ++i) {
            int64_t idx = base + (int64_t)threadIdx.x * kScanIPT + i;
            s += idx < n ? in[idx] : 0u;
        }
        uint32_t tot;
        block_exclusive_scan(s, ws, tot);  // ends with a barrier: ws is free again
        if (threadIdx.x == 0) partial[tile] = tot;
    }
}

// Exclusive scan of each tile, offset by the tile prefix: partial_scan[b]
// (already scanned), or with RAW the sum of the unscanned tile totals
// partial[0..b), which the workgroup adds up itself (few tiles: this saves
// the launches of scanning them).  in and out may alias.  If total !=
// null, the last thread of the last tile writes the grand total.
template <bool RAW>
__global__ void __launch_bounds__(kScanWG) k_scan_down(const uint32_t* in, int64_t n, const uint32_t* partial_scan,
                                                       uint32_t* out, uint32_t* total, const uint32_t* plan) {
    if (plan && plan_incremental(plan)) return;
    __shared__ uint32_t ws[kScanWG / 64];
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        uint32_t tile_prefix = 0;
        if (partial_scan) {
            if (RAW) {
                uint32_t part = 0;
                for (int64_t t = threadIdx.x; t < tile; t += kScanWG) part += partial_scan[t];
                block_exclusive_scan(part, ws, tile_prefix);
            } else {
                tile_prefix = partial_scan[tile];
            }
        }
        int64_t base = tile * kScanTile;
        uint32_t v[kScanIPT];
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < kScanIPT; ++i) {
            int64_t idx = base + (int64_t)threadIdx.x * kScanIPT + i;
            v[i] = idx < n ? in[idx] : 0u;
            s += v[i];
        }
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(s, ws, tot);
        uint32_t off = tile_prefix + pre;
#pragma unroll
        for (int i = 0; i < kScanIPT; ++i) {
            int64_t idx = base + (int64_t)threadIdx.x * kScanIPT + i;
            if (idx < n) out[idx] = off;
            off += v[i];
        }
        if (total && tile == tiles - 1 && threadIdx.x == kScanWG - 1) *total = off;
    }
}

int64_t scan_temp_words(int64_t n) {
    int64_t words = 0;
    while (n > kScanTile) {
        int64_t tiles = (n + kScanTile - 1) / kScanTile;
        words += tiles;
        n = tiles;
    }
    return words + 64;
}

// Workgroups of a launch over `tiles` tiles: one per tile, or at most
// max_wgs (> 0), each looping over its tiles.
static unsigned tile_grid(int64_t tiles, int max_wgs) {
    return (unsigned)(max_wgs > 0 && tiles > max_wgs ? max_wgs : tiles);
}

// Exclusive scan in place over n words; temp holds scan_temp_words(n).
// total (device, may be null) receives the sum.
void exclusive_scan(uint32_t* data, int64_t n, uint32_t* temp, uint32_t* total, hipStream_t st,
                    const uint32_t* plan, int max_wgs) {
    if (n <= 0) return;
    if (n <= kScanTile) {
        hipLaunchKernelGGL(k_scan_down<false>, dim3(1), dim3(kScanWG), 0, st, data, n, (const uint32_t*)nullptr, data,
                           total, plan);
        return;
    }
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    uint32_t* partial = temp;
    const unsigned grid = tile_grid(tiles, max_wgs);
    hipLaunchKernelGGL(k_scan_reduce, dim3(grid), dim3(kScanWG), 0, st, data, n, partial, plan);
    if (tiles <= 2048) {  // each tile sums at most 2047 tile totals itself
        hipLaunchKernelGGL(k_scan_down<true>, dim3(grid), dim3(kScanWG), 0, st, data, n, (const uint32_t*)partial,
                           data, total, plan);
        return;
    }
    exclusive_scan(partial, tiles, temp + tiles, nullptr, st, plan, max_wgs);
    hipLaunchKernelGGL(k_scan_down<false>, dim3(grid), dim3(kScanWG), 0, st, data, n, (const uint32_t*)partial, data,
                       total, plan);
}

int64_t radix_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }
// Sorts of at most this many tiles skip the separate digit-table scan.
constexpr int64_t kRawTiles = 32;
int64_t radix_hist_words(int64_t n) { return radix_tiles(n) * kMaxRadix; }

// Digit widths.  Large sorts (>= 2^20 keys) use digits of at most 7 bits:
// the scatter of a 2048-key tile then writes runs of >= 16 keys per digit,
// and on the 8.4M-particle slab (25-bit keys) 4 passes of 7/6/6/6 bits take
// 0.31 ms against 0.37 ms for 3 passes of 9/8/8 and 0.35 ms for 5 passes of
// 5-6 bits.  Small sorts are launch-bound and use as few passes as 8- or
// 9-bit digits allow, and so do large ones with few_passes (the caller
// expects the sort's launches to return at once most of the time: fewer
// launches, a slower sort when it does run).  Bits are split as evenly as
// possible, widest first.
int radix_passes(int key_bits, int64_t n, bool few_passes) {
    if (n >= (int64_t(1) << 20) && !few_passes) return (key_bits + 6) / 7;
    const int p8 = (key_bits + 7) / 8, p9 = (key_bits + 8) / 9;
    return p9 < p8 ? p9 : p8;
}
static int digit_bits(int key_bits, int64_t n, int pass, bool few_passes) {
    const int passes = radix_passes(key_bits, n, few_passes);
    int done = 0;
    for (int p = 0; p < pass; ++p) done += (key_bits - done + (passes - p) - 1) / (passes - p);
    return (key_bits - done + (passes - pass) - 1) / (passes - pass);
}

// Sorts (kin,vin) (read only) into (ka,va)/(kb,vb) ping-pong buffers;
// returns 0 if the result is in (ka,va), 1 if in (kb,vb).
int radix_sort_pairs(const uint32_t* kin, const uint32_t* vin, uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb,
                     int64_t n, int key_bits, uint32_t* hist, uint32_t* scan_temp, hipStream_t st,
                     const ProfileHook& prof, const uint32_t* plan, bool few_passes) {
    if (n <= 0) return 0;
    int64_t tiles = radix_tiles(n);
    int passes = radix_passes(key_bits, n, few_passes);
    // with few_passes the launches mostly return at once: fewer workgroups
    const int max_wgs = few_passes ? kLightGrid : 0;
    const unsigned grid = tile_grid(tiles, max_wgs);
    const uint32_t *ki = kin, *vi = vin;
    uint32_t *ko = ka, *vo = va;
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
        const int bits = digit_bits(key_bits, n, p, few_passes);
        const uint32_t mask = (1u << bits) - 1u;
        // digit widths of the plan: 9 or 8 bits, or 6-7 under a narrower plan
        const int dw = bits > 8 ? 9 : (bits > 7 ? 8 : (bits > 6 ? 7 : 6));
        const int radix = 1 << dw;
        prof.begin("sort_hist");
#define MPM_HIST(B) hipLaunchKernelGGL(k_radix_hist<B>, dim3(grid), dim3(kSortWG), 0, st, ki, n, shift, mask, hist, \
                                       tiles, plan)
        if (dw == 9) MPM_HIST(9);
        else if (dw == 8) MPM_HIST(8);
        else if (dw == 7) MPM_HIST(7);
        else MPM_HIST(6);
#undef MPM_HIST
        prof.end("sort_hist");
        // up to kRawTiles tiles the scatter reads the raw digit table itself
        // (one launch fewer per pass for small sorts)
        const bool raw = tiles <= kRawTiles;
        prof.begin("sort_scan");
        if (!raw) exclusive_scan(hist, tiles * radix, scan_temp, nullptr, st, plan, max_wgs);
        prof.end("sort_scan");
        prof.begin("sort_scatter");
#define MPM_SCATTER(B, R) hipLaunchKernelGGL((k_radix_scatter<B, R>), dim3(grid), dim3(kSortWG), 0, st, ki, vi, ko, \
                                             vo, n, shift, mask, hist, tiles, plan)
        if (raw) {
            if (dw == 9) MPM_SCATTER(9, true);
            else if (dw == 8) MPM_SCATTER(8, true);
            else if (dw == 7) MPM_SCATTER(7, true);
            else MPM_SCATTER(6, true);
        } else {
            if (dw == 9) MPM_SCATTER(9, false);
            else if (dw == 8) MPM_SCATTER(8, false);
            else if (dw == 7) MPM_SCATTER(7, false);
            else MPM_SCATTER(6, false);
        }
#undef MPM_SCATTER
        prof.end("sort_scatter");
        shift += bits;
        ki = ko;
        vi = vo;
        ko = (ko == ka) ? kb : ka;
        vo = (vo == va) ? vb : va;
    }
    return (passes & 1) ? 0 : 1;
}

}  // namespace mpm
