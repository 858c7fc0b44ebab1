// blockscan.h -- workgroup-wide exclusive scan of one uint32 per thread
// (wave64 shuffles, then the wave totals through LDS); shared by the
// device-wide scan and the work-item builder.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int kScanWG = 256;
constexpr int kScanIPT = 4;
constexpr int kScanTile = kScanWG * kScanIPT;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wave_sums, uint32_t& block_total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t up = __shfl_up(incl, off, 64);
        if (lane >= off) incl += up;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int w = 0; w < kScanWG / 64; ++w) {
        uint32_t s = wave_sums[w];
        if (w < wave) pre += s;
        tot += s;
    }
    __syncthreads();
    block_total = tot;
    return pre + incl - v;
}

}  // namespace mpm
