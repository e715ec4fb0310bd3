#pragma once

// Yarn/yarn self-contact on a uniform grid built by a device radix sort.
//
// The reference keeps a static all-pairs schedule (E(E-1)/2 pairs in
// colour batches, reference tools/cloth_bag_metal.mm:1154-1272) and walks the
// active batches in one threadgroup (reference src/metal/ClothBag.metal:
// 2125-2448); neither scales past a few thousand segments. Here the contact
// semantics are defined once, for the GPU and the CPU oracle alike:
//
// Broad phase, once per substep after the positions are predicted:
//   box(s) = [min, max] over both endpoints at the substep start (prev) and
//            at the predicted position, grown by e = yarn radius + margin.
//   candidates = every pair s < t whose boxes overlap (closed intervals on
//            all three axes) and whose segments are not topologically local
//            (no endpoint of t within two distance hops of an endpoint of s,
//            the locality rule of reference tools/cloth_bag_metal.mm:1107-1152).
// The grid only accelerates that set: cell = floor((v - origin) / h) with
// origin = the componentwise minimum box corner and h the configured cell
// size; each segment is entered once per cell its box covers, (cell id,
// segment) entries are sorted by a stable LSD radix sort, and a pair is
// reported only from its owner cell, the cell of max(lo_s, lo_t), so each
// pair appears exactly once, as (min, max). Entries and pairs are written
// through count / exclusive scan / emit passes, so the lists are identical
// on every run.
//
// The owner test needs no arithmetic: cell(v) = floor((v - origin) / h) is
// non-decreasing in v, so cell(max(lo_s, lo_t)) = max(cell(lo_s), cell(lo_t))
// per axis, and since both boxes cover the entry's cell c, that maximum is
// c exactly when cell(lo_s) = c or cell(lo_t) = c on the axis. Each entry
// carries those three "my box starts here" bits above the segment index.
//
// Narrow phase (reference solve_self_contact arithmetic, present and swept
// modes) runs on all candidates in parallel from one position snapshot
// (Jacobi). Each endpoint correction is converted to 2^-32 m fixed point and
// summed with 64-bit integer atomics, which are order independent; every
// touched knot then moves by the mean of its corrections. No float atomics.

#include "device_math.hpp"
#include "kernels_core.hpp"
#include "kernels_scan.hpp"

namespace numi_dev {

constexpr uint32_t kMaxCellsPerSegment = 512;
constexpr uint32_t kEntrySegmentBits = 29;               // entry value = segment | start flags << 29
constexpr uint32_t kEntrySegmentMask = (1u << kEntrySegmentBits) - 1u;
constexpr double kFixedScale = 4294967296.0;          // 2^32 per metre
constexpr double kFixedInverse = 1.0 / 4294967296.0;

struct SelfGrid {
    float4 origin;           // xyz origin of cell (0,0,0)
    float cell;
    uint32_t nx, ny, nz;
    uint32_t segments;       // D, scan length for the per-segment passes
    uint32_t entryCount;     // total entries requested by the boxes
    uint32_t sortCount;      // min(entryCount, capacity)
    uint32_t pairTotal;      // candidate pairs found
    uint32_t pairCount;      // min(pairTotal, capacity)
    uint32_t valid;          // 0 when the grid would exceed 2^32 cells
    uint32_t histLength;     // radix histogram length: kRadixBins * ceil(radixCount / kScanTile)
    uint32_t denseCells;     // nx * ny * nz when it fits the per-cell counters, else 0
    uint32_t cellScanLength; // denseCells when this substep sorts by counting, else 0
    uint32_t radixCount;     // sortCount when this substep sorts by radix, else 0
    uint32_t foundCount;     // candidates the run walks appended (may exceed the capacity)
    uint32_t pad[2];
};

// A sorted cell entry with its segment's box, the 32-byte record the
// candidate walk reads per visited entry (staged in LDS per run window):
// a = (cell id, segment | start flags << 29, lo.x, lo.y) as bits / floats,
// b = (lo.z, hi.x, hi.y, hi.z).
struct SortedEntry {
    float4 a, b;
};

// Per-knot sums of one Jacobi pass (self-contact corrections, point
// self-collision corrections or self-friction velocity changes), all order
// independent: the corrections and their lengths in 2^-32 fixed point and
// the longest length as float bits (integer max orders non-negative floats).
// A knot's bit in the pass's touched mask says it has contributions.
struct FixedAccumulator {
    long long x, y, z;
    long long lengthSum;
    uint32_t maxLength;
    uint32_t pad;
};

// ---------------------------------------------------------- broad phase --

__global__ void k_self_boxes(const NumiClothBagGPUConfig* __restrict__ cfg, float margin,
                             const float4* __restrict__ pos, const float4* __restrict__ prev,
                             const uint2* __restrict__ ij, uint32_t segments,
                             float4* __restrict__ boxLo, float4* __restrict__ boxHi,
                             float4* __restrict__ partialLo, float4* __restrict__ partialHi) {
    __shared__ float4 waveLo[kScanThreads / 64];
    __shared__ float4 waveHi[kScanThreads / 64];
    const float expand = cfg->clothMaterial.x + margin;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (s < segments) {
        const uint2 e = ij[s];
        const float4 p[4] = {pos[e.x], pos[e.y], prev[e.x], prev[e.y]};
        for (int k = 0; k < 4; ++k) {
            lo[0] = fminf(lo[0], p[k].x);
            lo[1] = fminf(lo[1], p[k].y);
            lo[2] = fminf(lo[2], p[k].z);
            hi[0] = fmaxf(hi[0], p[k].x);
            hi[1] = fmaxf(hi[1], p[k].y);
            hi[2] = fmaxf(hi[2], p[k].z);
        }
        for (int k = 0; k < 3; ++k) {
            lo[k] = lo[k] - expand;
            hi[k] = hi[k] + expand;
        }
        boxLo[s] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        boxHi[s] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    for (int k = 0; k < 3; ++k) {
        for (int offset = 32; offset > 0; offset >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], offset));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], offset));
        }
    }
    if ((threadIdx.x & 63u) == 0) {
        waveLo[threadIdx.x >> 6] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        waveHi[threadIdx.x >> 6] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 a = waveLo[0], b = waveHi[0];
        for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
            a = make_float4(fminf(a.x, waveLo[w].x), fminf(a.y, waveLo[w].y), fminf(a.z, waveLo[w].z), 0.0f);
            b = make_float4(fmaxf(b.x, waveHi[w].x), fmaxf(b.y, waveHi[w].y), fmaxf(b.z, waveHi[w].z), 0.0f);
        }
        partialLo[blockIdx.x] = a;
        partialHi[blockIdx.x] = b;
    }
}

__device__ __forceinline__ int cellOf(float v, float origin, float cell) {
    return static_cast<int>(floorf((v - origin) / cell));
}

// One workgroup: the grid origin and extent from the per-block box bounds
// (min/max are exact, so the reduction order does not matter).
__global__ void k_self_grid(const float4* __restrict__ partialLo, const float4* __restrict__ partialHi,
                            uint32_t blocks, const NumiClothBagGPUConfig* __restrict__ cfg,
                            const uint32_t* __restrict__ ok, uint32_t segments, uint32_t cellCapacity,
                            SelfGrid* __restrict__ grid, uint32_t* __restrict__ failure) {
    __shared__ float4 waveLo[kScanThreads / 64];
    __shared__ float4 waveHi[kScanThreads / 64];
    float l[3] = {INFINITY, INFINITY, INFINITY};
    float h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
        const float4 a = partialLo[b], c = partialHi[b];
        l[0] = fminf(l[0], a.x);
        l[1] = fminf(l[1], a.y);
        l[2] = fminf(l[2], a.z);
        h[0] = fmaxf(h[0], c.x);
        h[1] = fmaxf(h[1], c.y);
        h[2] = fmaxf(h[2], c.z);
    }
    for (int k = 0; k < 3; ++k)
        for (int offset = 32; offset > 0; offset >>= 1) {
            l[k] = fminf(l[k], __shfl_xor(l[k], offset));
            h[k] = fmaxf(h[k], __shfl_xor(h[k], offset));
        }
    if ((threadIdx.x & 63u) == 0) {
        waveLo[threadIdx.x >> 6] = make_float4(l[0], l[1], l[2], 0.0f);
        waveHi[threadIdx.x >> 6] = make_float4(h[0], h[1], h[2], 0.0f);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    float4 lo = waveLo[0], hi = waveHi[0];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
        lo = make_float4(fminf(lo.x, waveLo[w].x), fminf(lo.y, waveLo[w].y), fminf(lo.z, waveLo[w].z), 0.0f);
        hi = make_float4(fmaxf(hi.x, waveHi[w].x), fmaxf(hi.y, waveHi[w].y), fmaxf(hi.z, waveHi[w].z), 0.0f);
    }
    const float cell = cfg->clothMaterial.y;
    SelfGrid g{};
    g.origin = lo;
    g.cell = cell;
    g.segments = segments;
    bool valid = *ok != 0u && cell > 0.0f && isfinite(cell) && finite3(xyz(lo)) && finite3(xyz(hi));
    if (valid) {
        const double nx = static_cast<double>(cellOf(hi.x, lo.x, cell)) + 1.0;
        const double ny = static_cast<double>(cellOf(hi.y, lo.y, cell)) + 1.0;
        const double nz = static_cast<double>(cellOf(hi.z, lo.z, cell)) + 1.0;
        valid = nx * ny * nz < 4294967296.0;
        if (valid) {
            g.nx = static_cast<uint32_t>(nx);
            g.ny = static_cast<uint32_t>(ny);
            g.nz = static_cast<uint32_t>(nz);
            if (nx * ny * nz <= static_cast<double>(cellCapacity)) g.denseCells = g.nx * g.ny * g.nz;
        }
    }
    if (!valid && *ok != 0u) recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
    g.valid = valid ? 1u : 0u;
    *grid = g;
}

struct CellRange {
    int lo[3], hi[3];
    uint32_t count;
};

__device__ __forceinline__ CellRange cellRange(const float4& lo, const float4& hi, const SelfGrid& g) {
    CellRange r;
    r.lo[0] = cellOf(lo.x, g.origin.x, g.cell);
    r.lo[1] = cellOf(lo.y, g.origin.y, g.cell);
    r.lo[2] = cellOf(lo.z, g.origin.z, g.cell);
    r.hi[0] = cellOf(hi.x, g.origin.x, g.cell);
    r.hi[1] = cellOf(hi.y, g.origin.y, g.cell);
    r.hi[2] = cellOf(hi.z, g.origin.z, g.cell);
    const uint64_t n = static_cast<uint64_t>(r.hi[0] - r.lo[0] + 1) * (r.hi[1] - r.lo[1] + 1) *
                       (r.hi[2] - r.lo[2] + 1);
    r.count = n > kMaxCellsPerSegment ? 0u : static_cast<uint32_t>(n);
    return r;
}

__device__ __forceinline__ uint32_t cellId(int x, int y, int z, const SelfGrid& g) {
    return (static_cast<uint32_t>(z) * g.ny + static_cast<uint32_t>(y)) * g.nx + static_cast<uint32_t>(x);
}

// Per-segment entry counts; on a dense grid also clears the per-cell
// counters of the counting sort.
__global__ void k_self_cell_counts(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                   const SelfGrid* __restrict__ grid, uint32_t* __restrict__ counts,
                                   uint32_t* __restrict__ cellCounters, uint32_t* __restrict__ failure) {
    const SelfGrid g = *grid;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t c = s; c < g.denseCells; c += gridDim.x * blockDim.x) cellCounters[c] = 0u;
    if (s >= g.segments) return;
    uint32_t n = 0;
    if (g.valid) {
        const CellRange r = cellRange(boxLo[s], boxHi[s], g);
        n = r.count;
        if (n == 0) recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
    }
    counts[s] = n;
}

// The entries of segment s in emission order: its covered cells, x fastest,
// with the "box starts in this cell" bits of the owner test.
template <typename Visit>
__device__ __forceinline__ void forSegmentCells(uint32_t s, const CellRange& r, const SelfGrid& g, Visit visit) {
    for (int z = r.lo[2]; z <= r.hi[2]; ++z)
        for (int y = r.lo[1]; y <= r.hi[1]; ++y)
            for (int x = r.lo[0]; x <= r.hi[0]; ++x) {
                const uint32_t starts = (x == r.lo[0] ? 1u : 0u) | (y == r.lo[1] ? 2u : 0u) |
                                        (z == r.lo[2] ? 4u : 0u);
                visit(cellId(x, y, z, g), s | (starts << kEntrySegmentBits));
            }
}

// Entries are sorted by (cell, emission order) = (cell, segment), either by
// a counting sort over the grid's cells (dense grid, every entry fits) or by
// the LSD radix sort of the (cell, value) pairs written here in emission
// order. Every thread takes the same decision from the grid and the total.
// Both write the entries in emission order; the counting sort also keeps,
// per entry, the slot its cell counter's increment returned (its place
// among the cell's entries in an arbitrary order; the candidate pass ranks
// each cell's entries by segment).
__global__ void k_self_emit_entries(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                    SelfGrid* __restrict__ grid, const uint32_t* __restrict__ offsets,
                                    const uint32_t* __restrict__ entryTotal, uint32_t capacity,
                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                    uint32_t* __restrict__ cellCounters, uint32_t* __restrict__ cellSlot,
                                    uint32_t* __restrict__ failure) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const SelfGrid g = *grid;
    const uint32_t total = *entryTotal;
    const bool dense = g.valid && g.denseCells != 0u && total <= capacity;
    if (s == 0) {
        const uint32_t sorted = min(total, capacity);
        grid->entryCount = total;
        grid->sortCount = sorted;
        grid->radixCount = dense ? 0u : sorted;
        grid->cellScanLength = dense ? g.denseCells : 0u;
        grid->histLength = dense ? 0u : kRadixBins * ((sorted + kScanTile - 1) / kScanTile);
        if (total > capacity) recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
    }
    if (s >= g.segments || !g.valid) return;
    const CellRange r = cellRange(boxLo[s], boxHi[s], g);
    if (r.count == 0) return;
    uint32_t out = offsets[s];
    forSegmentCells(s, r, g, [&](uint32_t cell, uint32_t value) {
        if (out < capacity) {
            keys[out] = cell;
            vals[out] = value;
            if (dense) cellSlot[out] = atomicAdd(&cellCounters[cell], 1u);
        }
        ++out;
    });
}

// Counting sort placement: every entry to its cell's start (the exclusive
// scan of the cell counts) plus the slot it drew, so each cell's entries
// are contiguous and cells ascend (their order within a cell is the
// arbitrary slot order; the candidate pass ranks them by segment).
__global__ void k_self_place_entries(const SelfGrid* __restrict__ grid, const uint32_t* __restrict__ cellStarts,
                                     const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                     const uint32_t* __restrict__ cellSlot, uint32_t* __restrict__ keysOut,
                                     uint32_t* __restrict__ valsOut) {
    const SelfGrid g = *grid;
    if (g.cellScanLength == 0u) return;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < g.sortCount; k += gridDim.x * blockDim.x) {
        const uint32_t cell = keysIn[k];
        const uint32_t at = cellStarts[cell] + cellSlot[k];
        keysOut[at] = cell;
        valsOut[at] = valsIn[k];
    }
}

// Final position of each placed entry in the sorted order (two-hop lists
// searched in memory): its cell's start plus the number of entries of the
// cell with a smaller segment (a segment enters a cell at most once, so
// ranks are distinct). This is the order a stable sort by cell of the
// emission-ordered entries gives.
__global__ void k_self_rank_entries(const SelfGrid* __restrict__ grid, const uint32_t* __restrict__ cellStarts,
                                    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    const SelfGrid g = *grid;
    if (g.cellScanLength == 0u) return;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < g.sortCount; k += gridDim.x * blockDim.x) {
        const uint32_t cell = keysIn[k];
        const uint32_t v = valsIn[k];
        const uint32_t s = v & kEntrySegmentMask;
        const uint32_t begin = cellStarts[cell];
        const uint32_t end = cell + 1u < g.cellScanLength ? cellStarts[cell + 1u] : g.sortCount;
        uint32_t at = begin;
        for (uint32_t j = begin; j < end; ++j) at += (valsIn[j] & kEntrySegmentMask) < s ? 1u : 0u;
        keysOut[at] = cell;
        valsOut[at] = v;
    }
}

// Two knots are local when one lies in the other's two-hop neighbourhood
// (sorted CSR lists that include the knot itself).
__device__ __forceinline__ bool inList(const uint32_t* __restrict__ list, uint32_t begin, uint32_t end,
                                       uint32_t value) {
    while (begin < end) {
        const uint32_t mid = begin + (end - begin) / 2;
        const uint32_t v = list[mid];
        if (v == value) return true;
        if (v < value) begin = mid + 1;
        else end = mid;
    }
    return false;
}

__device__ __forceinline__ bool localPair(uint2 a, uint2 b, const uint32_t* __restrict__ start,
                                          const uint32_t* __restrict__ list) {
    const uint32_t a0 = start[a.x], a1 = start[a.x + 1];
    const uint32_t b0 = start[a.y], b1 = start[a.y + 1];
    return inList(list, a0, a1, b.x) || inList(list, a0, a1, b.y) || inList(list, b0, b1, b.x) ||
           inList(list, b0, b1, b.y);
}

__device__ __forceinline__ bool boxesOverlap(const float4& aLo, const float4& aHi, const float4& bLo,
                                             const float4& bHi) {
    return aLo.x <= bHi.x && aLo.y <= bHi.y && aLo.z <= bHi.z && aHi.x >= bLo.x && aHi.y >= bLo.y &&
           aHi.z >= bLo.z;
}

// Candidates of sorted entry k = (cell, s): the later entries (cell, t) of
// the same cell run, kept when the boxes overlap, this cell is the pair's
// owner cell (the cell of max(lo_s, lo_t)) and the segments are not local.
// Each pair is therefore visited exactly once, from one entry, and comes out
// as (min, max). Runs are short (the segments whose grown boxes cover one
// cell), so the work is spread over all entries without searching.
template <typename Emit>
__device__ void forEntryCandidates(uint32_t k, const SelfGrid& g, const float4* __restrict__ boxLo,
                                   const float4* __restrict__ boxHi, const uint2* __restrict__ ij,
                                   const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                   const uint32_t* __restrict__ localStart, const uint32_t* __restrict__ localList,
                                   Emit emit) {
    const uint32_t n = g.sortCount;
    const uint32_t id = keys[k];
    const uint32_t vs = vals[k];
    const uint32_t s = vs & kEntrySegmentMask;
    const float4 lo = boxLo[s], hi = boxHi[s];
    const uint2 es = ij[s];
    for (uint32_t j = k + 1; j < n && keys[j] == id; ++j) {
        const uint32_t vt = vals[j];
        if (((vs | vt) >> kEntrySegmentBits) != 7u) continue;  // not the owner cell
        const uint32_t t = vt & kEntrySegmentMask;
        const float4 tLo = boxLo[t], tHi = boxHi[t];
        if (!boxesOverlap(lo, hi, tLo, tHi)) continue;
        if (localPair(es, ij[t], localStart, localList)) continue;
        emit(min(s, t), max(s, t));
    }
}

// Whether knot x or y is in either two-hop list: one compare per slot and
// knot, the results OR-ed as wave masks (scalar instructions) and read back
// at this lane's bit.
template <uint32_t W>
__device__ __forceinline__ bool inLists(const uint32_t (&na)[W], const uint32_t (&nb)[W], uint32_t x, uint32_t y) {
    unsigned long long hits = 0;
#pragma unroll
    for (uint32_t m = 0; m < W; ++m)
        hits |= __ballot(na[m] == x) | __ballot(na[m] == y) | __ballot(nb[m] == x) | __ballot(nb[m] == y);
    return ((hits >> __lane_id()) & 1ull) != 0ull;
}

// The cyclic half of a cell run that sorted entry i walks: with the run at
// [r0, r1) (L entries) and i at offset q, the h(q) entries after i, wrapping
// from r1 - 1 to r0, with h = (L - 1) / 2, plus one for the first L / 2
// entries when L is even. Every unordered pair of the run lies in exactly
// one of the two entries' halves (their offsets differ by d and L - d, and
// exactly one of those is below L / 2, or both equal L / 2 and only the
// earlier entry takes that step), and every entry of a run walks the same
// number of records up to one, so the lanes of a wave, which hold
// consecutive entries, stay together (walking the later entries of the run
// instead leaves half of each wave idle on average).
__device__ __forceinline__ uint32_t runHalf(uint32_t q, uint32_t L) {
    return (L - 1u) / 2u + ((L & 1u) == 0u && q < L / 2u ? 1u : 0u);
}

// Candidates of sorted entry i over its cyclic half run (the pairs of a cell
// run are the same set as over the later entries of the run, so every
// candidate is still found exactly once, from its owner cell, as (min, max)).
// The two-hop knot lists of both endpoints (W slots each, padded with ~0u)
// are requested before the walk and tested only afterwards: the walk over
// the run's records (owner cell, box overlap, shared knot) marks the
// partners left for the locality test in a 64-bit mask, so the lists'
// latency overlaps it and every lane requests them once, together.
// entryAt(j) / ijAt(j) return sorted entry j's record and its segment's
// endpoints (from an LDS copy or global memory). For every 64 steps with
// candidates, reserve(n) is called with their number, then emit(p, q) for
// each in walk order (increasing step), so the list is identical on every
// run.
template <uint32_t W, typename EntryAt, typename IJAt, typename Reserve, typename Emit>
__device__ __forceinline__ void forRunCandidates(uint32_t i, uint32_t r0, uint32_t r1, EntryAt entryAt, IJAt ijAt,
                                                 const uint32_t* __restrict__ localFixed, Reserve reserve,
                                                 Emit emit) {
    const SortedEntry self = entryAt(i);
    const uint32_t vs = __float_as_uint(self.a.y);
    const uint32_t s = vs & kEntrySegmentMask;
    const float4 lo = make_float4(self.a.z, self.a.w, self.b.x, 0.0f);
    const float4 hi = make_float4(self.b.y, self.b.z, self.b.w, 0.0f);
    const uint2 es = ijAt(i);
    const uint32_t L = r1 - r0;
    const uint32_t h = runHalf(i - r0, L);
    if (h == 0u) return;
    uint32_t na[W], nb[W];
    {
        const uint4* ra = reinterpret_cast<const uint4*>(localFixed + static_cast<size_t>(es.x) * W);
        const uint4* rb = reinterpret_cast<const uint4*>(localFixed + static_cast<size_t>(es.y) * W);
#pragma unroll
        for (uint32_t m = 0; m < W / 4; ++m) {
            const uint4 va = ra[m], vb = rb[m];
            na[4 * m] = va.x;
            na[4 * m + 1] = va.y;
            na[4 * m + 2] = va.z;
            na[4 * m + 3] = va.w;
            nb[4 * m] = vb.x;
            nb[4 * m + 1] = vb.y;
            nb[4 * m + 2] = vb.z;
            nb[4 * m + 3] = vb.w;
        }
    }
    constexpr uint32_t kUnroll = 4;
    for (uint32_t d0 = 0; d0 < h; d0 += 64u) {
        const uint32_t dn = min(64u, h - d0);
        uint64_t pending = 0;
        // j = the partner at step d0 + d + 1 (i + d0 + 1 < r1 + L, so one
        // wrap suffices).
        uint32_t j = i + d0 + 1u;
        if (j >= r1) j -= L;
        for (uint32_t d = 0; d < dn; d += kUnroll) {
            SortedEntry e[kUnroll];
            uint2 et[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                if (d + u < dn) {
                    e[u] = entryAt(j);
                    et[u] = ijAt(j);
                }
                if (++j == r1) j = r0;
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
                if (d + u >= dn) continue;
                const uint32_t vt = __float_as_uint(e[u].a.y);
                if (((vs | vt) >> kEntrySegmentBits) != 7u) continue;  // not the owner cell
                const float4 tLo = make_float4(e[u].a.z, e[u].a.w, e[u].b.x, 0.0f);
                const float4 tHi = make_float4(e[u].b.y, e[u].b.z, e[u].b.w, 0.0f);
                if (!boxesOverlap(lo, hi, tLo, tHi)) continue;
                // A shared knot is in both two-hop lists (each list holds its
                // own knot): local without reading them.
                if (et[u].x == es.x || et[u].x == es.y || et[u].y == es.x || et[u].y == es.y) continue;
                pending |= 1ull << (d + u);
            }
        }
        uint64_t found = 0;
        for (uint64_t left = pending; left != 0ull; left &= left - 1ull) {
            const uint32_t d = static_cast<uint32_t>(__builtin_ctzll(left));
            uint32_t jt = i + d0 + d + 1u;
            if (jt >= r1) jt -= L;
            const uint2 et = ijAt(jt);
            if (!inLists<W>(na, nb, et.x, et.y)) found |= 1ull << d;
        }
        if (found == 0ull) continue;
        reserve(static_cast<uint32_t>(__popcll(found)));
        for (; found != 0ull; found &= found - 1ull) {
            const uint32_t d = static_cast<uint32_t>(__builtin_ctzll(found));
            uint32_t jt = i + d0 + d + 1u;
            if (jt >= r1) jt -= L;
            const uint32_t t = __float_as_uint(entryAt(jt).a.y) & kEntrySegmentMask;
            emit(min(s, t), max(s, t));
        }
    }
}

struct PairArgs {
    const float4* boxLo;
    const float4* boxHi;
    const uint2* ij;
    SelfGrid* grid;
    // The sorted entries (cell id, segment | start flags): the radix sort
    // leaves them here; the counting sort's candidate pass writes them here
    // from the placed entries, whose cells are contiguous but unranked.
    uint32_t* keys;
    uint32_t* vals;
    const uint32_t* placedKeys;
    const uint32_t* placedVals;
    const uint32_t* localStart;
    const uint32_t* localList;
    const uint32_t* localFixed;
    uint32_t* counts;
    const uint32_t* offsets;
    const uint32_t* pairTotal;
    uint32_t capacity;
    uint2* pairs;
    uint4* pairKnots;  // the candidates' four endpoint knots (s's two, then t's), for the solves
    NumiClothBagGPUSelfImpulse* impulses;  // per-candidate impulse aggregates, cleared here
    uint32_t* failure;
    // Run walks (W > 0): the candidates the count pass found, appended as
    // (entry, rank among the entry's candidates, s, t) in any order (the
    // emission pass puts each at its entry's scanned offset plus its rank),
    // and the list's capacity (the candidates' capacity).
    uint4* found;
};

// Sorted entry j's record, read from the sorted entries and its segment's box.
__device__ __forceinline__ SortedEntry sortedRecord(const PairArgs& a, uint32_t j) {
    const uint32_t v = a.vals[j];
    const uint32_t s = v & kEntrySegmentMask;
    const float4 lo = a.boxLo[s], hi = a.boxHi[s];
    SortedEntry e;
    e.a = make_float4(__uint_as_float(a.keys[j]), __uint_as_float(v), lo.x, lo.y);
    e.b = make_float4(lo.z, hi.x, hi.y, hi.z);
    return e;
}

constexpr uint32_t kPairThreads = 256;  // workgroup size of the candidate count passes

// Emits a candidate at out (the entry's next scanned offset) and clears its
// impulse aggregate; slots past the capacity are dropped (a RANGE failure).
__device__ __forceinline__ void emitCandidate(const PairArgs& a, uint32_t& out, uint32_t p, uint32_t q) {
    if (out < a.capacity) {
        a.pairs[out] = make_uint2(p, q);
        const uint2 ep = a.ij[p], eq = a.ij[q];
        a.pairKnots[out] = make_uint4(ep.x, ep.y, eq.x, eq.y);
        NumiClothBagGPUSelfImpulse empty{};
        empty.identity = mr_uint4{out, p, q, 0u};
        a.impulses[out] = empty;
    }
    ++out;
}

__device__ __forceinline__ void publishPairTotal(const PairArgs& a, const SelfGrid& g) {
    const uint32_t total = g.valid ? *a.pairTotal : 0u;
    a.grid->pairTotal = total;
    a.grid->pairCount = min(total, a.capacity);
    if (total > a.capacity) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_RANGE);
}

// Two-hop lists searched in memory (meshes whose lists exceed 32 knots):
// one thread per entry, walking the later entries of its run.
__global__ void __launch_bounds__(kPairThreads) k_self_pair_counts(PairArgs a) {
    const SelfGrid g = *a.grid;
    if (!g.valid) return;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < g.sortCount; k += gridDim.x * blockDim.x) {
        uint32_t n = 0;
        forEntryCandidates(k, g, a.boxLo, a.boxHi, a.ij, a.keys, a.vals, a.localStart, a.localList,
                           [&](uint32_t, uint32_t) { ++n; });
        a.counts[k] = n;
    }
}

__global__ void k_self_emit_pairs(PairArgs a) {
    const SelfGrid g = *a.grid;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) publishPairTotal(a, g);
    if (!g.valid) return;
    for (uint32_t e = k; e < g.sortCount; e += gridDim.x * blockDim.x) {
        uint32_t out = a.offsets[e];
        // The offsets are the exclusive scan of the counts: an entry with no
        // candidate (almost all of them) has nothing to emit.
        const uint32_t end = e + 1 < g.sortCount ? a.offsets[e + 1] : *a.pairTotal;
        if (end == out) continue;
        forEntryCandidates(e, g, a.boxLo, a.boxHi, a.ij, a.keys, a.vals, a.localStart, a.localList,
                           [&](uint32_t p, uint32_t q) { emitCandidate(a, out, p, q); });
    }
}

// Run walks with the two-hop lists in registers (W = 16 or 32 slots).
// A workgroup owns the cell runs that start in a chunk of Chunk sorted
// entries and stages the window of Window entries from the chunk's start
// in LDS, which holds every owned run unless a run is longer than the
// window's tail; such a run is ranked and walked in global memory. Chunks
// of 640 in windows of 768 keep four workgroups on a CU (a window of 2048
// for the bag's runs of up to ~130 entries in 10 cm cells measured slower:
// one workgroup per CU).
constexpr uint32_t kRunChunk = 640;
constexpr uint32_t kRunWindow = 768;

// The end of the run of key `key` that continues past `from` (the keys of
// the sorted or placed entries), found by the whole workgroup 256 entries at
// a time.
__device__ uint32_t runEndFrom(const uint32_t* __restrict__ keys, uint32_t n, uint32_t from, uint32_t key,
                               uint32_t* found) {
    for (uint32_t base = from;; base += kPairThreads) {
        if (threadIdx.x == 0) *found = ~0u;
        __syncthreads();
        const uint32_t j = base + threadIdx.x;
        if (j >= n || keys[j] != key) atomicMin(found, j);
        __syncthreads();
        const uint32_t end = *found;
        __syncthreads();
        if (end != ~0u) return min(end, n);
    }
}

// Candidate counts of every entry of the owned runs, and the candidates
// themselves appended to the found list. The input is the placed entries
// of the counting sort (each cell's entries contiguous, in slot order) or
// the radix sort's sorted entries: each entry of an owned run is ranked by
// segment within its run (a segment enters a cell once, so ranks are
// distinct; for sorted input the rank is the entry's own place), its record
// (cell, value, box, endpoints) goes to LDS at that place, and the counting
// sort's sorted entries are written out, so the walk sees the order of a
// stable sort by cell whichever sort ran.
template <uint32_t W, uint32_t Chunk, uint32_t Window>
__global__ void __launch_bounds__(kPairThreads) k_self_run_counts(PairArgs a) {
    constexpr uint32_t kSlots = Window / kPairThreads;  // window slots per thread, contiguous
    static_assert(kSlots * kPairThreads == Window && Chunk < Window && Window <= 65535u, "run window layout");
    // First the window's keys and values, then the ranked records.
    __shared__ SortedEntry rec[Window];
    __shared__ uint2 recIJ[Window];
    __shared__ uint32_t runStart[Window];  // absolute start of the window's k-th run
    __shared__ uint16_t runOf[Window];     // run index of a window slot (0xffff: run began before the chunk)
    __shared__ uint32_t scratch[kPairThreads / 64];
    __shared__ uint32_t info[3];  // runs in the window, (unused), run end search
    static_assert(2 * sizeof(uint32_t) * Window <= sizeof(rec), "window keys and values alias the records");
    uint32_t* const winKey = reinterpret_cast<uint32_t*>(rec);
    uint32_t* const winVal = winKey + Window;
    const SelfGrid g = *a.grid;
    if (!g.valid) return;
    const uint32_t n = g.sortCount;
    const bool counting = g.cellScanLength != 0u;
    const uint32_t* __restrict__ keysIn = counting ? a.placedKeys : a.keys;
    const uint32_t* __restrict__ valsIn = counting ? a.placedVals : a.vals;
    const uint32_t t = threadIdx.x;
    for (uint32_t c0 = blockIdx.x * Chunk; c0 < n; c0 += gridDim.x * Chunk) {
        const uint32_t wn = min(Window, n - c0);
        const uint32_t k0 = t * kSlots;
        uint32_t keys[kSlots], vals[kSlots];
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) {
            const uint32_t k = k0 + u;
            keys[u] = k < wn ? keysIn[c0 + k] : 0u;
            vals[u] = k < wn ? valsIn[c0 + k] : 0u;
            if (k < wn) {
                winKey[k] = keys[u];
                winVal[k] = vals[u];
            }
        }
        const uint32_t keyBefore = c0 == 0u ? ~0u : keysIn[c0 - 1u];
        __syncthreads();
        // Run starts of the window: a block-wide scan of the start flags
        // numbers the runs and gives every slot its run.
        uint32_t flags = 0, count = 0;
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) {
            const uint32_t k = k0 + u;
            if (k >= wn) continue;
            const uint32_t before = k == 0u ? keyBefore : (u == 0u ? winKey[k - 1u] : keys[u - 1u]);
            if (keys[u] != before) {
                flags |= 1u << u;
                ++count;
            }
        }
        uint32_t r = blockExclusiveScan256(count, scratch);
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) {
            const uint32_t k = k0 + u;
            if (k >= wn) continue;
            if ((flags >> u) & 1u) runStart[r++] = c0 + k;
            runOf[k] = r == 0u ? static_cast<uint16_t>(0xffffu) : static_cast<uint16_t>(r - 1u);
        }
        if (t == kPairThreads - 1u) info[0] = r;
        __syncthreads();
        const uint32_t runs = info[0];
        // Owned runs start in the chunk's first Chunk slots.
        uint32_t owned = runs;
        if (wn > Chunk) owned = runOf[Chunk - 1u] == 0xffffu ? 0u : runOf[Chunk - 1u] + 1u;
        if (owned == 0u) {
            __syncthreads();
            continue;
        }
        const uint32_t first = runStart[0];
        uint32_t ownedEnd;
        if (owned < runs) {
            ownedEnd = runStart[owned];
        } else if (c0 + wn == n) {
            ownedEnd = n;
        } else {
            // The last owned run reaches past the window (a run longer than
            // the window's tail): find its end in global memory.
            ownedEnd = runEndFrom(keysIn, n, c0 + wn, winKey[runStart[owned - 1u] - c0], &info[2]);
        }
        // Entries from `giant` on belong to a run that reaches past the
        // window; it is ranked and walked in global memory.
        const uint32_t giant = ownedEnd > c0 + wn ? runStart[owned - 1u] : ownedEnd;
        uint32_t place[kSlots];
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) {
            const uint32_t k = k0 + u;
            const uint32_t i = c0 + k;
            place[u] = ~0u;
            if (k >= wn || i < first || i >= giant) continue;
            const uint32_t run = runOf[k];
            const uint32_t r0 = runStart[run];
            const uint32_t r1 = run + 1u < runs ? runStart[run + 1u] : ownedEnd;
            const uint32_t seg = vals[u] & kEntrySegmentMask;
            uint32_t rank = 0;
            for (uint32_t j = r0 - c0; j < r1 - c0; ++j) rank += (winVal[j] & kEntrySegmentMask) < seg ? 1u : 0u;
            place[u] = r0 + rank;
        }
        __syncthreads();  // the window's keys and values are read no more
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) {
            if (place[u] == ~0u) continue;
            const uint32_t seg = vals[u] & kEntrySegmentMask;
            const float4 lo = a.boxLo[seg], hi = a.boxHi[seg];
            SortedEntry e;
            e.a = make_float4(__uint_as_float(keys[u]), __uint_as_float(vals[u]), lo.x, lo.y);
            e.b = make_float4(lo.z, hi.x, hi.y, hi.z);
            rec[place[u] - c0] = e;
            recIJ[place[u] - c0] = a.ij[seg];
            if (counting) {
                a.keys[place[u]] = keys[u];
                a.vals[place[u]] = vals[u];
            }
        }
        if (giant < ownedEnd && counting) {
            for (uint32_t i = giant + t; i < ownedEnd; i += kPairThreads) {
                const uint32_t v = valsIn[i];
                const uint32_t seg = v & kEntrySegmentMask;
                uint32_t rank = 0;
                for (uint32_t j = giant; j < ownedEnd; ++j) rank += (valsIn[j] & kEntrySegmentMask) < seg ? 1u : 0u;
                a.keys[giant + rank] = keysIn[i];
                a.vals[giant + rank] = v;
            }
        }
        __syncthreads();
        const auto ldsEntry = [&](uint32_t j) { return rec[j - c0]; };
        const auto ldsIJ = [&](uint32_t j) { return recIJ[j - c0]; };
        const auto globalEntry = [&](uint32_t j) { return sortedRecord(a, j); };
        const auto globalIJ = [&](uint32_t j) { return a.ij[a.vals[j] & kEntrySegmentMask]; };
        for (uint32_t i = first + t; i < ownedEnd; i += kPairThreads) {
            uint32_t r0, r1;
            if (i >= giant) {
                r0 = giant;
                r1 = ownedEnd;
            } else {
                const uint32_t run = runOf[i - c0];
                r0 = runStart[run];
                r1 = run + 1u < runs ? runStart[run + 1u] : ownedEnd;
            }
            // Each group of candidates takes slots of the found list; every
            // record carries its rank among the entry's candidates.
            uint32_t total = 0, slot = 0;
            const auto reserve = [&](uint32_t m) {
                slot = atomicAdd(&a.grid->foundCount, m);
            };
            const auto emit = [&](uint32_t p, uint32_t q) {
                if (slot < a.capacity) a.found[slot] = make_uint4(i, total, p, q);
                ++slot;
                ++total;
            };
            if (i < giant) forRunCandidates<W>(i, r0, r1, ldsEntry, ldsIJ, a.localFixed, reserve, emit);
            else forRunCandidates<W>(i, r0, r1, globalEntry, globalIJ, a.localFixed, reserve, emit);
            a.counts[i] = total;
        }
        __syncthreads();
    }
}

// Candidates at their entries' scanned offsets plus their ranks: a scatter
// of the found list. Should the list have overflowed (more candidates than
// their capacity, a RANGE failure, and then which ones the list kept
// depends on timing), every entry with candidates is walked again, its run
// bounds read from the keys, and the first `capacity` candidates of the
// scanned order are written, the same list on every run.
template <uint32_t W>
__global__ void __launch_bounds__(kScanThreads) k_self_emit_run_pairs(PairArgs a) {
    const SelfGrid g = *a.grid;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) publishPairTotal(a, g);
    if (!g.valid) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (g.foundCount <= a.capacity) {
        for (uint32_t u = k; u < g.foundCount; u += stride) {
            const uint4 c = a.found[u];
            uint32_t out = a.offsets[c.x] + c.y;
            emitCandidate(a, out, c.z, c.w);
        }
        return;
    }
    const auto entryAt = [&](uint32_t j) { return sortedRecord(a, j); };
    const auto ijAt = [&](uint32_t j) { return a.ij[a.vals[j] & kEntrySegmentMask]; };
    const uint32_t n = g.sortCount;
    for (uint32_t e = k; e < n; e += stride) {
        uint32_t out = a.offsets[e];
        const uint32_t end = e + 1 < n ? a.offsets[e + 1] : *a.pairTotal;
        if (end == out) continue;
        const uint32_t key = a.keys[e];
        uint32_t r0 = e, r1 = e + 1u;
        while (r0 > 0u && a.keys[r0 - 1u] == key) --r0;
        while (r1 < n && a.keys[r1] == key) ++r1;
        forRunCandidates<W>(e, r0, r1, entryAt, ijAt, a.localFixed, [](uint32_t) {},
                            [&](uint32_t p, uint32_t q) { emitCandidate(a, out, p, q); });
    }
}

// ---------------------------------------------------------- narrow phase --

struct SegmentSample {
    V3 first, second;
    float firstWeight, secondWeight, distance;
};

// Closest points of two segments (reference sampleSegments,
// src/metal/ClothBag.metal:135-211).
__device__ SegmentSample sampleSegments(V3 a0, V3 a1, V3 b0, V3 b1) {
    const V3 da = a1 - a0;
    const V3 db = b1 - b0;
    const V3 offset = a0 - b0;
    const float la = dot(da, da);
    const float lb = dot(db, db);
    const float pb = dot(db, offset);
    float wa = 0.0f, wb = 0.0f;
    if (la <= 1.0e-20f && lb <= 1.0e-20f) return {a0, b0, 0.0f, 0.0f, length(b0 - a0)};
    if (la <= 1.0e-20f) {
        wb = clampf(pb / lb, 0.0f, 1.0f);
    } else {
        const float pa = dot(da, offset);
        if (lb <= 1.0e-20f) {
            wa = clampf(-pa / la, 0.0f, 1.0f);
        } else {
            const float cross_ = dot(da, db);
            const float denom = la * lb - cross_ * cross_;
            if (denom > 1.0e-20f) wa = clampf((cross_ * pb - pa * lb) / denom, 0.0f, 1.0f);
            wb = (cross_ * wa + pb) / lb;
            if (wb < 0.0f) {
                wb = 0.0f;
                wa = clampf(-pa / la, 0.0f, 1.0f);
            } else if (wb > 1.0f) {
                wb = 1.0f;
                wa = clampf((cross_ - pa) / la, 0.0f, 1.0f);
            }
        }
    }
    const V3 pA = fma3(da, wa, a0);
    const V3 pB = fma3(db, wb, b0);
    return {pA, pB, wa, wb, length(pB - pA)};
}

__device__ V3 contactNormal(V3 separation, V3 da, V3 db, V3 previousOffset) {
    const float len = length(separation);
    if (len > 1.0e-12f) return separation / len;
    V3 n = safeNormalized(cross(da, db));
    if (dot(previousOffset, n) > 0.0f) n = -n;
    return n;
}

// The move of a touched knot: the sum of its corrections scaled by
// (longest length) / (sum of lengths). One correction moves the knot by
// itself, n equal ones by one of them (their mean), and a correction that
// shrinks to zero drops out continuously; a mean by contribution count would
// jump whenever a contact at its threshold appears or vanishes, and a knot
// pressed by one deep and several grazing contacts would keep most of its
// penetration. Never longer than the longest correction.
__device__ __forceinline__ V3 fixedMean(const FixedAccumulator& a) {
    if (a.lengthSum <= 0) return v3(0.0f, 0.0f, 0.0f);
    const double scale = static_cast<double>(__uint_as_float(a.maxLength)) / static_cast<double>(a.lengthSum);
    return v3(static_cast<float>(static_cast<double>(a.x) * scale), static_cast<float>(static_cast<double>(a.y) * scale),
              static_cast<float>(static_cast<double>(a.z) * scale));
}

// Adds d to knot i's fixed-point sum and sets the knot's bit in the touched
// mask, so the apply pass visits only the knots that have corrections.
__device__ __forceinline__ void addFixed(FixedAccumulator* acc, uint32_t* touched, uint32_t i, V3 d) {
    FixedAccumulator* a = acc + i;
    atomicAdd(reinterpret_cast<unsigned long long*>(&a->x),
              static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(d.x) * kFixedScale)));
    atomicAdd(reinterpret_cast<unsigned long long*>(&a->y),
              static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(d.y) * kFixedScale)));
    atomicAdd(reinterpret_cast<unsigned long long*>(&a->z),
              static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(d.z) * kFixedScale)));
    const float len = length(d);
    atomicAdd(reinterpret_cast<unsigned long long*>(&a->lengthSum),
              static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(len) * kFixedScale)));
    atomicMax(&a->maxLength, __float_as_uint(len));
    // Every contribution sets the bit (idempotent; no atomic waits for a
    // returned value).
    atomicOr(touched + (i >> 5), 1u << (i & 31u));
}

// solve_self_contact (reference src/metal/ClothBag.metal:2125-2448) for every
// candidate against one snapshot; mode 0 present overlap, mode 1 swept
// conservative advancement (80 steps, 0.9 safety, 1e-10 minimum advance).
// pairKnots: each candidate's endpoint knots, written with the candidate (one
// dependent load per candidate instead of the pair, then its segments).
// The solve of k_self_solve by workgroup `block` of `blocks` (the fruit/yarn
// solve runs it as the one workgroup of its launch).
__device__ __forceinline__ void selfSolve(const NumiClothBagGPUConfig* __restrict__ cfg, uint32_t mode,
                                          const SelfGrid* __restrict__ grid, const uint4* __restrict__ pairKnots,
                                          const float4* __restrict__ pos, const float4* __restrict__ prev,
                                          FixedAccumulator* __restrict__ acc, uint32_t* __restrict__ touched,
                                          NumiClothBagGPUSelfImpulse* __restrict__ impulses,
                                          NumiClothBagGPUSelfStatus* __restrict__ status, uint32_t block,
                                          uint32_t blocks) {
    const uint32_t count = grid->valid ? grid->pairCount : 0u;
    const float r = cfg->clothMaterial.x;
    const float target = 2.0f * r;
    const float tolerance = 1.0e-9f;
    const uint32_t stride = blocks * blockDim.x;
    const uint32_t rounds = (count + stride - 1) / stride;
    for (uint32_t round = 0; round < rounds; ++round) {
        const uint32_t k = round * stride + block * blockDim.x + threadIdx.x;
        bool accepted = false;
        float correction = 0.0f;
        if (k < count) {
            const uint4 kk = pairKnots[k];
            const uint2 ea = make_uint2(kk.x, kk.y);
            const uint2 eb = make_uint2(kk.z, kk.w);
            const float4 A0 = pos[ea.x], A1 = pos[ea.y], B0 = pos[eb.x], B1 = pos[eb.y];
            // The start positions: the swept mode reads them always, the
            // present mode only for a contact normal of coincident points.
            float4 Q0{}, Q1{}, R0{}, R1{};
            if (mode == 1u) {
                Q0 = prev[ea.x];
                Q1 = prev[ea.y];
                R0 = prev[eb.x];
                R1 = prev[eb.y];
            }
            // edgeBoundsOverlap (reference :213-274) with expansion r.
            float aLo[3] = {fminf(A0.x, A1.x), fminf(A0.y, A1.y), fminf(A0.z, A1.z)};
            float aHi[3] = {fmaxf(A0.x, A1.x), fmaxf(A0.y, A1.y), fmaxf(A0.z, A1.z)};
            float bLo[3] = {fminf(B0.x, B1.x), fminf(B0.y, B1.y), fminf(B0.z, B1.z)};
            float bHi[3] = {fmaxf(B0.x, B1.x), fmaxf(B0.y, B1.y), fmaxf(B0.z, B1.z)};
            if (mode == 1u) {
                const float qa[2][3] = {{Q0.x, Q0.y, Q0.z}, {Q1.x, Q1.y, Q1.z}};
                const float qb[2][3] = {{R0.x, R0.y, R0.z}, {R1.x, R1.y, R1.z}};
                for (int a = 0; a < 3; ++a) {
                    aLo[a] = fminf(aLo[a], fminf(qa[0][a], qa[1][a]));
                    aHi[a] = fmaxf(aHi[a], fmaxf(qa[0][a], qa[1][a]));
                    bLo[a] = fminf(bLo[a], fminf(qb[0][a], qb[1][a]));
                    bHi[a] = fmaxf(bHi[a], fmaxf(qb[0][a], qb[1][a]));
                }
            }
            bool overlap = true;
            for (int a = 0; a < 3; ++a) overlap = overlap && (aLo[a] - r <= bHi[a] + r) && (aHi[a] + r >= bLo[a] - r);
            float wa = 0.0f, wb = 0.0f;
            V3 n = v3(0.0f, 0.0f, 0.0f);
            if (overlap) {
                if (mode == 0u) {
                    const SegmentSample c = sampleSegments(xyz(A0), xyz(A1), xyz(B0), xyz(B1));
                    if (c.distance < target) {
                        wa = c.firstWeight;
                        wb = c.secondWeight;
                        const V3 separation = c.second - c.first;
                        V3 previousOffset = v3(0.0f, 0.0f, 0.0f);  // read by contactNormal only at zero separation
                        if (!(length(separation) > 1.0e-12f)) {
                            const V3 pa = mix3(xyz(prev[ea.x]), xyz(prev[ea.y]), wa);
                            const V3 pb = mix3(xyz(prev[eb.x]), xyz(prev[eb.y]), wb);
                            previousOffset = pa - pb;
                        }
                        n = contactNormal(separation, xyz(A1) - xyz(A0), xyz(B1) - xyz(B0), previousOffset);
                        correction = target - c.distance;
                    }
                } else {
                    const float bound = length(xyz(A0) - xyz(Q0)) + length(xyz(A1) - xyz(Q1)) +
                                        length(xyz(B0) - xyz(R0)) + length(xyz(B1) - xyz(R1));
                    if (bound >= 1.0e-14f) {
                        float t = 0.0f;
                        auto at = [&](float time) {
                            return sampleSegments(mix3(xyz(Q0), xyz(A0), time), mix3(xyz(Q1), xyz(A1), time),
                                                  mix3(xyz(R0), xyz(B0), time), mix3(xyz(R1), xyz(B1), time));
                        };
                        SegmentSample hit = at(0.0f);
                        bool found = hit.distance <= target + tolerance;
                        if (!found) {
                            for (int step = 0; step < 80; ++step) {
                                hit = at(t);
                                const float gap = hit.distance - target;
                                if (gap <= tolerance) {
                                    found = true;
                                    break;
                                }
                                const float advance = 0.9f * gap / bound;
                                if (!isfinite(advance) || !(advance > 0.0f) || t + advance >= 1.0f) break;
                                t += fmaxf(advance, 1.0e-10f);
                            }
                        }
                        if (found) {
                            wa = hit.firstWeight;
                            wb = hit.secondWeight;
                            const V3 pa = mix3(xyz(Q0), xyz(Q1), wa);
                            const V3 pb = mix3(xyz(R0), xyz(R1), wb);
                            const V3 a0t = mix3(xyz(Q0), xyz(A0), t), a1t = mix3(xyz(Q1), xyz(A1), t);
                            const V3 b0t = mix3(xyz(R0), xyz(B0), t), b1t = mix3(xyz(R1), xyz(B1), t);
                            n = contactNormal(hit.second - hit.first, a1t - a0t, b1t - b0t, pa - pb);
                            const V3 restA = (xyz(A0) - a0t) * (1.0f - wa) + (xyz(A1) - a1t) * wa;
                            const V3 restB = (xyz(B0) - b0t) * (1.0f - wb) + (xyz(B1) - b1t) * wb;
                            correction = dot(restA - restB, n);
                            if (!(correction > 0.0f)) correction = 0.0f;
                        }
                    }
                }
            }
            if (correction > 0.0f) {
                const float ua = 1.0f - wa, ub = 1.0f - wb;
                const float denom = A0.w * ua * ua + A1.w * wa * wa + B0.w * ub * ub + B1.w * wb * wb;
                if (denom > 0.0f && isfinite(denom)) {
                    const float lambda = correction / denom;
                    addFixed(acc, touched, ea.x, n * -(A0.w * ua * lambda));
                    addFixed(acc, touched, ea.y, n * -(A1.w * wa * lambda));
                    addFixed(acc, touched, eb.x, n * (B0.w * ub * lambda));
                    addFixed(acc, touched, eb.y, n * (B1.w * wb * lambda));
                    accepted = true;
                    // The contact's impulse record (reference :2391-2407),
                    // summed into the candidate's aggregate in call order
                    // (reference aggregate_self_impulses :2503-2553). Only
                    // this thread touches candidate k in a call.
                    const float impulse = lambda / cfg->gravityAndTimestep.w;
                    NumiClothBagGPUSelfImpulse a = impulses[k];
                    a.identity.w += 1u;
                    a.normalAndImpulse.x = a.normalAndImpulse.x + -n.x * impulse;
                    a.normalAndImpulse.y = a.normalAndImpulse.y + -n.y * impulse;
                    a.normalAndImpulse.z = a.normalAndImpulse.z + -n.z * impulse;
                    a.normalAndImpulse.w = a.normalAndImpulse.w + impulse;
                    a.endpointImpulses.x = a.endpointImpulses.x + ua * impulse;
                    a.endpointImpulses.y = a.endpointImpulses.y + wa * impulse;
                    a.endpointImpulses.z = a.endpointImpulses.z + ub * impulse;
                    a.endpointImpulses.w = a.endpointImpulses.w + wb * impulse;
                    impulses[k] = a;
                }
            }
        }
        const unsigned long long ballot = __ballot(accepted);
        float m = accepted ? correction : 0.0f;
        for (int offset = 32; offset > 0; offset >>= 1) m = fmaxf(m, __shfl_xor(m, offset));
        if ((threadIdx.x & 63u) == 0 && ballot != 0ull) {
            atomicAdd(mode == 0u ? &status->counters.x : &status->counters.y, static_cast<uint32_t>(__popcll(ballot)));
            atomicMax(&status->counters.z, __float_as_uint(m));
        }
    }
}

__global__ void k_self_solve(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                             uint32_t mode, const SelfGrid* __restrict__ grid, const uint4* __restrict__ pairKnots,
                             const float4* __restrict__ pos,
                             const float4* __restrict__ prev, FixedAccumulator* __restrict__ acc,
                             uint32_t* __restrict__ touched, NumiClothBagGPUSelfImpulse* __restrict__ impulses,
                             NumiClothBagGPUSelfStatus* __restrict__ status) {
    if (*ok == 0u) return;
    selfSolve(cfg, mode, grid, pairKnots, pos, prev, acc, touched, impulses, status, blockIdx.x, gridDim.x);
}

// Moves every touched knot by the mean of its fixed-point corrections and
// clears its accumulator and mask bit for the next call. One thread per
// 32-knot word of the touched mask.
// pos = nullptr: clear only (a buffer whose corrections were applied by the
// tile programs that loaded the knots).
__device__ __forceinline__ void applyPendingCorrection(float4& p, uint32_t i, uint32_t word, bool live,
                                                       FixedAccumulator* __restrict__ acc,
                                                       uint32_t* __restrict__ failure) {
    (void)word;
    if (live) {
        const V3 np = xyz(p) + fixedMean(acc[i]);
        if (!finite3(np)) recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        else p = withXyz(p, np);
    }
    acc[i] = FixedAccumulator{0, 0, 0, 0, 0u, 0u};
}

__global__ void k_self_apply(const uint32_t* __restrict__ ok, float4* __restrict__ pos,
                             FixedAccumulator* __restrict__ acc, uint32_t* __restrict__ touched, uint32_t words,
                             uint32_t* __restrict__ failure) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t bits = touched[w];
    if (bits == 0u) return;
    touched[w] = 0u;
    while (bits != 0u) {
        const uint32_t i = (w << 5) + static_cast<uint32_t>(__ffs(bits) - 1);
        bits &= bits - 1u;
        if (pos == nullptr || *ok == 0u) {
            acc[i] = FixedAccumulator{0, 0, 0, 0, 0u, 0u};
            continue;
        }
        const V3 d = fixedMean(acc[i]);
        const float4 p = pos[i];
        const V3 np = xyz(p) + d;
        if (!finite3(np)) recordFailure(failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        else pos[i] = withXyz(p, np);
        acc[i] = FixedAccumulator{0, 0, 0, 0, 0u, 0u};
    }
}

// Self-contact friction (reference apply_self_friction,
// src/metal/ClothBag.metal:2580-2731) on the candidates that made contact in
// the substep: the aggregate's normal, normal impulse and endpoint weights
// give a maximum-dissipation tangential impulse limited by mu times the
// normal impulse. The reference walks a static pair colouring in order; the
// hash candidates have none, so every pair acts on one velocity snapshot
// (Jacobi) and each touched knot takes the mean of its velocity changes,
// summed in 2^-32 m/s fixed point (order independent), as the normal
// response does with positions. Contact count -> metrics.w, largest
// impulse / limit -> metrics.x (integer max on float bits).
__global__ void k_self_friction(const NumiClothBagGPUConfig* __restrict__ cfg, const uint32_t* __restrict__ ok,
                                const SelfGrid* __restrict__ grid, const uint2* __restrict__ pairs,
                                const uint2* __restrict__ ij, const float4* __restrict__ pos,
                                const float4* __restrict__ vel, const NumiClothBagGPUSelfImpulse* __restrict__ impulses,
                                FixedAccumulator* __restrict__ acc, uint32_t* __restrict__ touched,
                                NumiClothBagGPUFrictionStatus* __restrict__ friction) {
    if (*ok == 0u) return;
    const uint32_t count = grid->valid ? grid->pairCount : 0u;
    const float mu = cfg->clothMaterial.w;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (count + stride - 1) / stride;
    for (uint32_t round = 0; round < rounds; ++round) {
        const uint32_t k = round * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool applied = false;
        float ratio = 0.0f;
        if (k < count && impulses[k].identity.w != 0u) {
            const NumiClothBagGPUSelfImpulse c = impulses[k];
            const uint2 pr = pairs[k];
            const uint2 ea = ij[pr.x];
            const uint2 eb = ij[pr.y];
            const float normalImpulse = c.normalAndImpulse.w;
            const V3 nsum = v3(c.normalAndImpulse.x, c.normalAndImpulse.y, c.normalAndImpulse.z);
            const float normalLength = length(nsum);
            const float firstSum = c.endpointImpulses.x + c.endpointImpulses.y;
            const float secondSum = c.endpointImpulses.z + c.endpointImpulses.w;
            if (normalImpulse > 0.0f && normalLength > 1.0e-10f && firstSum > 1.0e-12f && secondSum > 1.0e-12f) {
                const V3 normal = nsum / normalLength;
                const float f0 = c.endpointImpulses.x / firstSum, f1 = c.endpointImpulses.y / firstSum;
                const float s0 = c.endpointImpulses.z / secondSum, s1 = c.endpointImpulses.w / secondSum;
                const float4 A0 = pos[ea.x], A1 = pos[ea.y], B0 = pos[eb.x], B1 = pos[eb.y];
                const V3 firstVelocity = xyz(vel[ea.x]) * f0 + xyz(vel[ea.y]) * f1;
                const V3 secondVelocity = xyz(vel[eb.x]) * s0 + xyz(vel[eb.y]) * s1;
                const V3 relative = firstVelocity - secondVelocity;
                const V3 tangentVelocity = relative - normal * dot(relative, normal);
                const float slip = length(tangentVelocity);
                if (slip > 1.0e-10f) {
                    const V3 tangent = tangentVelocity / slip;
                    const float denominator = A0.w * f0 * f0 + A1.w * f1 * f1 + B0.w * s0 * s0 + B1.w * s1 * s1;
                    if (denominator > 0.0f) {
                        const float limit = mu * normalImpulse;
                        const float tangential = fminf(slip / denominator, limit);
                        if (tangential > 0.0f) {
                            const V3 onFirst = tangent * -tangential;
                            addFixed(acc, touched, ea.x, onFirst * (A0.w * f0));
                            addFixed(acc, touched, ea.y, onFirst * (A1.w * f1));
                            addFixed(acc, touched, eb.x, -(onFirst * (B0.w * s0)));
                            addFixed(acc, touched, eb.y, -(onFirst * (B1.w * s1)));
                            applied = true;
                            if (limit > 0.0f) ratio = tangential / limit;
                        }
                    }
                }
            }
        }
        const unsigned long long ballot = __ballot(applied);
        for (int offset = 32; offset > 0; offset >>= 1) ratio = fmaxf(ratio, __shfl_xor(ratio, offset));
        if ((threadIdx.x & 63u) == 0 && ballot != 0ull) {
            atomicAdd(&friction->metrics.w, static_cast<uint32_t>(__popcll(ballot)));
            atomicMax(&friction->metrics.x, __float_as_uint(ratio));
        }
    }
}

// Largest capsule overlap 2r - distance over the current candidates
// (reference maximumSelfPenetration, tools/cloth_bag_metal.mm:5216-5276),
// per-block partial maxima written into the w lane of the diagnostic
// partials.
__global__ void k_self_penetration_into(const SelfGrid* __restrict__ grid, const uint2* __restrict__ pairs,
                                        const uint2* __restrict__ ij, const float4* __restrict__ pos, float radius,
                                        float4* __restrict__ partial) {
    __shared__ float waveMaxima[kScanThreads / 64];
    const uint32_t count = grid->valid ? grid->pairCount : 0u;
    float m = 0.0f;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const uint2 pr = pairs[k];
        const uint2 a = ij[pr.x], b = ij[pr.y];
        const SegmentSample c = sampleSegments(xyz(pos[a.x]), xyz(pos[a.y]), xyz(pos[b.x]), xyz(pos[b.y]));
        m = fmaxf(m, 2.0f * radius - c.distance);
    }
    for (int offset = 32; offset > 0; offset >>= 1) m = fmaxf(m, __shfl_xor(m, offset));
    if ((threadIdx.x & 63u) == 0) waveMaxima[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) r = fmaxf(r, waveMaxima[w]);
        partial[blockIdx.x].w = r;
    }
}

}  // namespace numi_dev
