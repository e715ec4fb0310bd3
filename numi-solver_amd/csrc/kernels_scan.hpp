#pragma once

// Device exclusive scan of uint32 arrays and a stable LSD radix sort of
// (uint32 key, uint32 value) pairs, both deterministic (integer arithmetic
// only, fixed reduction order). Element counts live in device memory so the
// launches can sit inside a captured graph; each workgroup walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... up to the live count, so a grid of
// at most kScanBlocks workgroups serves any capacity and a launch over few
// live elements costs few workgroups.

#include "device_math.hpp"

namespace numi_dev {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 2048
constexpr uint32_t kRadixBits = 8;
constexpr uint32_t kRadixBins = 1u << kRadixBits;
constexpr uint32_t kScanBlocks = 2048;

// Exclusive scan of the 256 thread totals held in LDS; returns the block total.
__device__ uint32_t blockExclusiveScan256(uint32_t value, uint32_t* scratch) {
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    // Inclusive scan within the wave (Hillis-Steele over 64 lanes).
    uint32_t inclusive = value;
    for (uint32_t offset = 1; offset < 64; offset <<= 1) {
        const uint32_t other = __shfl_up(inclusive, offset);
        if (lane >= offset) inclusive += other;
    }
    if (lane == 63) scratch[wave] = inclusive;
    __syncthreads();
    uint32_t waveBase = 0;
    uint32_t total = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) waveBase += scratch[w];
        total += scratch[w];
    }
    __syncthreads();
    (void)total;
    return inclusive - value + waveBase;  // exclusive
}

__device__ uint32_t blockTotal256(uint32_t value, uint32_t* scratch) {
    uint32_t v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset);
    if ((threadIdx.x & 63u) == 0) scratch[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t total = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) total += scratch[w];
    __syncthreads();
    return total;
}

// Phase 1: per-tile exclusive scan in place, tile totals to `sums`.
__global__ void k_scan_tiles(uint32_t* __restrict__ data, const uint32_t* __restrict__ count,
                             uint32_t* __restrict__ sums) {
    __shared__ uint32_t scratch[kScanThreads / 64];
    const uint32_t n = *count;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        uint32_t items[kScanItems];
        uint32_t local = 0;
        const uint32_t first = tile * kScanTile + threadIdx.x * kScanItems;
        for (uint32_t k = 0; k < kScanItems; ++k) {
            const uint32_t i = first + k;
            items[k] = i < n ? data[i] : 0u;
            local += items[k];
        }
        const uint32_t offset = blockExclusiveScan256(local, scratch);
        const uint32_t total = blockTotal256(local, scratch);
        uint32_t running = offset;
        for (uint32_t k = 0; k < kScanItems; ++k) {
            const uint32_t i = first + k;
            if (i < n) data[i] = running;
            running += items[k];
        }
        if (threadIdx.x == 0) sums[tile] = total;
    }
}

// Phase 2: one workgroup scans the tile totals sequentially in chunks of 256
// and stores the grand total.
__global__ void k_scan_sums(uint32_t* __restrict__ sums, const uint32_t* __restrict__ count,
                            uint32_t* __restrict__ total) {
    __shared__ uint32_t scratch[kScanThreads / 64];
    const uint32_t tiles = (*count + kScanTile - 1) / kScanTile;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < tiles; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < tiles ? sums[i] : 0u;
        const uint32_t offset = blockExclusiveScan256(v, scratch);
        const uint32_t chunk = blockTotal256(v, scratch);
        if (i < tiles) sums[i] = carry + offset;
        carry += chunk;
    }
    if (threadIdx.x == 0) *total = carry;
}

// Phase 3: add the scanned tile offsets.
__global__ void k_scan_add(uint32_t* __restrict__ data, const uint32_t* __restrict__ count,
                           const uint32_t* __restrict__ sums) {
    const uint32_t n = *count;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t add = sums[tile];
        const uint32_t base = tile * kScanTile;
        for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanThreads) {
            const uint32_t i = base + k;
            if (i < n) data[i] += add;
        }
    }
}

// ---------------------------------------------------------------- radix --

// Per-tile digit histogram, stored digit-major over the live tiles
// (ceil(*count / kScanTile)): hist[digit * tiles + tile], so one exclusive
// scan of kRadixBins * tiles entries yields every (digit, tile) output base
// in stable order.
__global__ void k_radix_histogram(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ count,
                                  uint32_t shift, uint32_t* __restrict__ hist) {
    __shared__ uint32_t bins[kRadixBins];
    const uint32_t n = *count;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        bins[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t base = tile * kScanTile;
        for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanThreads) {
            const uint32_t i = base + k;
            if (i < n) atomicAdd(&bins[(keys[i] >> shift) & (kRadixBins - 1)], 1u);
        }
        __syncthreads();
        hist[threadIdx.x * tiles + tile] = bins[threadIdx.x];
        __syncthreads();
    }
}

// Stable scatter. The tile is consumed in 8 rounds of 256 consecutive items;
// within a round each lane's rank among equal digits comes from ballots
// (8 digit bits -> same-digit lane mask), waves are ordered through LDS, and
// per-digit running counters carry across rounds.
__global__ void k_radix_scatter(const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                const uint32_t* __restrict__ count, uint32_t shift,
                                const uint32_t* __restrict__ histScanned, uint32_t* __restrict__ keysOut,
                                uint32_t* __restrict__ valsOut) {
    __shared__ uint32_t running[kRadixBins];
    __shared__ uint32_t waveCount[kScanThreads / 64][kRadixBins];
    const uint32_t n = *count;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t base = tile * kScanTile;
        running[tid] = histScanned[tid * tiles + tile];
        for (uint32_t w = 0; w < kScanThreads / 64; ++w) waveCount[w][tid] = 0;
        __syncthreads();
        for (uint32_t round = 0; round < kScanItems; ++round) {
            const uint32_t i = base + round * kScanThreads + tid;
            const bool live = i < n;
            const uint32_t key = live ? keysIn[i] : 0u;
            const uint32_t value = live ? valsIn[i] : 0u;
            const uint32_t digit = (key >> shift) & (kRadixBins - 1);
            unsigned long long same = __ballot(live);
            for (uint32_t b = 0; b < kRadixBits; ++b) {
                const unsigned long long ones = __ballot(live && ((digit >> b) & 1u));
                same &= ((digit >> b) & 1u) ? ones : ~ones;
            }
            const uint32_t rank = static_cast<uint32_t>(__popcll(same & below));
            if (live && rank == 0) waveCount[wave][digit] = static_cast<uint32_t>(__popcll(same));
            __syncthreads();
            if (live) {
                uint32_t position = running[digit] + rank;
                for (uint32_t w = 0; w < wave; ++w) position += waveCount[w][digit];
                keysOut[position] = key;
                valsOut[position] = value;
            }
            __syncthreads();
            uint32_t add = 0;
            for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
                add += waveCount[w][tid];
                waveCount[w][tid] = 0;
            }
            running[tid] += add;
            __syncthreads();
        }
    }
}

}  // namespace numi_dev
