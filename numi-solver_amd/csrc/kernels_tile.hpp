#pragma once

// LDS-resident tile programs (see tile_plan.hpp for the plan and its
// exactness argument). One workgroup per tile:
//   1. gather the tile's loaded knots (owned + halo) from posIn into LDS,
//   1b. optionally apply the self-contact corrections of the previous
//      self-contact solve and the ground/sphere pass that followed it
//      (both act on one knot at a time, so halo copies stay exact),
//   2. run every pass of the program on LDS, one item per thread slot, with a
//      workgroup barrier between colours; multipliers are read from the
//      input buffer and written to the output buffer by the owning tile only,
//   3. run the trailing per-knot ground/sphere passes,
//   4. write the owned knots to posOut.
// Inputs and outputs are separate buffers (ping-pong across launches), so a
// tile never reads a value another tile of the same launch has written.
// The per-constraint arithmetic is the same device function the per-colour
// kernels call, so results are bit-identical to the per-colour schedule.

#include "kernels_core.hpp"
#include "kernels_self.hpp"

namespace numi_dev {

constexpr uint32_t kMaxTilePasses = 64;
constexpr uint32_t kXcds = 8;           // accelerator complex dies of an MI355X
constexpr uint32_t kTileStaticLds = 2048;  // bound on k_tile_program's static LDS (checked below)
constexpr uint32_t kTileDistanceKind = 0;  // values of numi_hip::TilePassKind
constexpr uint32_t kTileKnotKind = 1;
constexpr uint32_t kTileBendKind = 2;
constexpr uint32_t kTileStrainKind = 3;

struct TileProgramArgs {
    const NumiClothBagGPUConfig* cfg;
    const uint32_t* ok;
    const float4* posIn;
    float4* posOut;
    const uint32_t* loadStart;
    const uint32_t* loadIds;
    const uint32_t* ownedCount;
    const uint32_t* itemStart;
    const uint4* items;         // 16-byte items; strain programs: two 8-byte items per uint4
    const float2* distanceRestCompliance;
    const float* distanceMaxLength;
    const float* distanceLambdaIn;
    float* distanceLambdaOut;
    const float2* bendChordCompliance;
    const float* bendLambdaIn;
    float* bendLambdaOut;
    const float2* knotCosineCompliance;
    const float* knotLambdaIn;
    float* knotLambdaOut;
    const float4* spheres;
    uint32_t sphereCount;
    uint32_t sphereWindowMin;   // lists of at least this many spheres use a sphere window
    uint32_t passes;
    uint32_t sweepPasses;       // strain programs: passes per sweep (0 otherwise)
    // Self-contact corrections to apply first (null: none), and the buffer of
    // an earlier solve to clear for reuse (null: none). Other tiles read the
    // applied buffer in this launch, so it is cleared by a later one.
    const FixedAccumulator* selfAcc;
    const uint32_t* selfTouched;
    FixedAccumulator* clearAcc;
    uint32_t* clearTouched;
    uint32_t leadingCollide;    // 1: the ground/sphere pass before the first colour pass
    uint32_t trailingCollides;
    // Staged programs: LDS offset (float4 units) of the staged items and the
    // item capacity; the staged input multipliers follow the items.
    uint32_t stageOffset;
    uint32_t stageItems;
    uint32_t* failure;
    // Strain programs of the reconciliation rounds: tileChanged (null: not
    // tracked) receives, per tile, whether the tile's owned knots leave the
    // launch different from how they entered it. With prevChanged (the
    // previous round's flags) a tile returns at once when no tile owning one
    // of its loaded knots (tileNbrs[tileNbrStart[tile]..], itself included)
    // changed them in the previous round. The host sets prevChanged only on a
    // program that repeats the previous round's with nothing in between and
    // no leading pass: the tile's loaded knots are then the previous round's
    // output, equal to that round's input for these tiles, so its owned
    // knots would come out as that round produced them, which is what the
    // output buffer (that round's input) already holds.
    uint32_t* tileChanged;
    const uint32_t* prevChanged;
    const uint32_t* tileNbrStart;
    const uint32_t* tileNbrs;
    unsigned long long* roundStats;  // tracked tile-rounds: [0] skipped, [1] run
    // Sphere windows built once per substep by k_tile_windows (null: each
    // ground/sphere pass builds its own): per tile the window and its
    // sphereCount-slot list. A window holds for any knot inside its box
    // (knots that leave it are redone with the whole list), so one built
    // from the loaded knots at the substep's start serves every pass.
    const SphereWindow* tileWindows;
    const float4* tileWindowSpheres;
    uint32_t kinds[kMaxTilePasses];
};

// Ground/sphere pass over every loaded knot (owned and halo: the pass acts on
// one knot at a time, so halo knots stay equal to their owners' values).
// Non-finite knots are skipped and reported by their owning tile only.
// windowReady: the window in LDS is the tile's substep window, loaded by an
// earlier pass of the launch.
__device__ __forceinline__ void tileCollide(float4* lds, uint32_t tile, uint32_t loaded, uint32_t owned,
                                            const TileProgramArgs& a, bool ground, float radius,
                                            const float4* sphereCache, float4* culled, SphereWindow& window,
                                            float4* scratch, bool& windowReady) {
    const bool windowed = a.sphereCount >= a.sphereWindowMin;
    if (windowed && a.tileWindows != nullptr) {
        if (!windowReady) {
            if (threadIdx.x == 0) window = a.tileWindows[tile];
            for (uint32_t k = threadIdx.x; k < a.sphereCount; k += blockDim.x)
                culled[k] = a.tileWindowSpheres[static_cast<size_t>(tile) * a.sphereCount + k];
            __syncthreads();
            windowReady = true;
        }
    } else if (windowed) {
        V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t i = threadIdx.x; i < loaded; i += blockDim.x) {
            const float4 p = lds[i];
            if (!isfinite(p.z)) continue;
            lo = v3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
            hi = v3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
        }
        buildSphereWindow(lo, hi, radius, sphereCache, a.sphereCount, culled, window, scratch);
    }
    for (uint32_t i = threadIdx.x; i < loaded; i += blockDim.x) {
        float4 p = lds[i];
        if (!isfinite(p.z)) {
            if (i < owned) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
            continue;
        }
        const bool moved = windowed
                               ? collideKnotWindow(p, ground, radius, window, culled, sphereCache, a.sphereCount)
                               : collideKnot(p, ground, radius, sphereCache, a.sphereCount);
        if (moved) lds[i] = p;
    }
    __syncthreads();
}

constexpr uint32_t kTilePrefetch = 4;  // item slots per thread and pass held in registers
static_assert(kTilePrefetch % 2 == 0, "register slots are processed in pairs");
constexpr uint32_t kTileKnotSlots = 4; // loaded knots per thread whose ids stay in registers

// Ids of a thread's first kTileKnotSlots loaded knots (load list entries
// t + u * stride; slots past the end repeat the last entry so every slot
// holds a valid knot id).
__device__ __forceinline__ void tileLoadKnotIds(const TileProgramArgs& a, uint32_t base, uint32_t loaded, uint32_t t,
                                                uint32_t stride, uint32_t* ids) {
    const uint32_t last = loaded > 0u ? loaded - 1u : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kTileKnotSlots; ++u)
        ids[u] = loaded > 0u ? a.loadIds[base + min(t + u * stride, last)] : 0u;
}

// The tile's knots into LDS (no barrier), each with the self-contact
// correction of the previous solve applied (mean of its fixed-point sum,
// live runs only), and the owned knots' entries of the accumulator buffer
// being recycled cleared, one mask bit at a time (other tiles clear the
// other bits of a shared word). For the register slots every request is in
// flight before any is used: the position and both mask words together,
// then the accumulators of touched knots; inFlight() runs in between
// (further independent loads of the caller).
// orig receives the register slots' positions as read.
template <typename InFlight>
__device__ __forceinline__ void tileGatherKnots(float4* lds, const TileProgramArgs& a, uint32_t base, uint32_t loaded,
                                                uint32_t owned, uint32_t t, uint32_t stride, bool live,
                                                const uint32_t* ids, float4 (&orig)[kTileKnotSlots],
                                                InFlight&& inFlight) {
    const bool applySelf = live && a.selfAcc != nullptr;
    const bool clearing = a.clearAcc != nullptr;
    float4 p[kTileKnotSlots];
    uint32_t selfWord[kTileKnotSlots], clearWord[kTileKnotSlots];
#pragma unroll
    for (uint32_t u = 0; u < kTileKnotSlots; ++u) {
        p[u] = a.posIn[ids[u]];
        selfWord[u] = applySelf ? a.selfTouched[ids[u] >> 5] : 0u;
        clearWord[u] = clearing ? a.clearTouched[ids[u] >> 5] : 0u;
    }
    inFlight();
#pragma unroll
    for (uint32_t u = 0; u < kTileKnotSlots; ++u) orig[u] = p[u];
#pragma unroll
    for (uint32_t u = 0; u < kTileKnotSlots; ++u) {
        const uint32_t i = t + u * stride;
        if (i >= loaded) break;
        const uint32_t id = ids[u];
        const uint32_t bit = 1u << (id & 31u);
        if ((selfWord[u] & bit) != 0u) {
            const V3 np = xyz(p[u]) + fixedMean(a.selfAcc[id]);
            if (!finite3(np)) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
            else p[u] = withXyz(p[u], np);
        }
        lds[i] = p[u];
        if (i < owned && (clearWord[u] & bit) != 0u) {
            a.clearAcc[id] = FixedAccumulator{0, 0, 0, 0, 0u, 0u};
            atomicAnd(a.clearTouched + (id >> 5), ~bit);
        }
    }
    for (uint32_t i = t + kTileKnotSlots * stride; i < loaded; i += stride) {
        const uint32_t id = a.loadIds[base + i];
        const uint32_t bit = 1u << (id & 31u);
        float4 q = a.posIn[id];
        if (applySelf && (a.selfTouched[id >> 5] & bit) != 0u) {
            const V3 np = xyz(q) + fixedMean(a.selfAcc[id]);
            if (!finite3(np)) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
            else q = withXyz(q, np);
        }
        lds[i] = q;
        if (clearing && i < owned && (a.clearTouched[id >> 5] & bit) != 0u) {
            a.clearAcc[id] = FixedAccumulator{0, 0, 0, 0, 0u, 0u};
            atomicAnd(a.clearTouched + (id >> 5), ~bit);
        }
    }
}

__device__ __forceinline__ uint32_t tileItemConstraint(uint32_t kind, const uint4& item) {
    return (kind == kTileKnotKind ? item.z : item.y) & 0x7fffffffu;
}

__device__ __forceinline__ bool tileItemOwner(uint32_t kind, const uint4& item) {
    return ((kind == kTileKnotKind ? item.z : item.y) >> 31) != 0u && kind != kTileStrainKind;
}

// Item j of a program. Strain programs store 8-byte items, two per uint4
// (x = local a | local b << 16, y = rest * (1 + limit) as bits): the strain
// pass reads nothing else, and its passes are bound by these loads. The
// item is returned in x and y of a uint4 either way, so every kernel keeps
// the same register arrays.
template <bool Strain>
__device__ __forceinline__ uint4 tileItemAt(const uint4* __restrict__ items, uint32_t j) {
    if constexpr (Strain) {
        const uint2 v = reinterpret_cast<const uint2*>(items)[j];
        return make_uint4(v.x, v.y, 0u, 0u);
    } else {
        return items[j];
    }
}

// Item slots m = 0..kTilePrefetch-1 of thread t in pass k (zero past the end).
template <bool Strain>
__device__ __forceinline__ void tileLoadItems(const uint4* __restrict__ items, const uint32_t* __restrict__ itemStart,
                                              uint32_t passes, uint32_t k, uint32_t t, uint32_t stride, uint4* dst) {
    if (k >= passes) return;
    const uint32_t first = itemStart[k], last = itemStart[k + 1];
#pragma unroll
    for (uint32_t m = 0; m < kTilePrefetch; ++m) {
        const uint32_t j = first + t + m * stride;
        dst[m] = j < last ? tileItemAt<Strain>(items, j) : make_uint4(0u, 0u, 0u, 0u);
    }
}

// The input multipliers of those slots (in: the pass kind's input buffer).
__device__ __forceinline__ void tileLoadLambdas(const float* __restrict__ in, const uint32_t* __restrict__ itemStart,
                                                uint32_t passes, uint32_t k, uint32_t kind, uint32_t t,
                                                uint32_t stride, const uint4* src, float* dst) {
    if (k >= passes) return;
    const uint32_t first = itemStart[k], last = itemStart[k + 1];
#pragma unroll
    for (uint32_t m = 0; m < kTilePrefetch; ++m)
        dst[m] = first + t + m * stride < last ? in[tileItemConstraint(kind, src[m])] : 0.0f;
}

// One constraint of a pass on the LDS knots; lam is its multiplier from the
// input buffer, written to the output buffer when this tile owns it.
// Strain = true handles strain-limit items only and returns whether the item
// moved its knots; Strain = false handles distance, knot and bend items.
// out is the pass kind's output multiplier buffer.
template <bool Strain>
__device__ __forceinline__ bool tileItem(float4* lds, uint32_t kind, const uint4& item, float lam, float* out,
                                         const float2* __restrict__ knotCC, uint32_t* failure, bool live, float dt,
                                         float invDt2, bool ground, float floorZ) {
    if constexpr (Strain) {
        const uint32_t la = item.x & 0xffffu, lb = item.x >> 16;
        if (!live) return false;
        float4 p = lds[la], q = lds[lb];
        if (!limitStrainPair(p, q, __uint_as_float(item.y), failure)) return false;
        lds[la] = p;
        lds[lb] = q;
        return true;
    }
    const uint32_t la = item.x & 0xffffu, lb = item.x >> 16;
    const uint32_t c = tileItemConstraint(kind, item);
    if (kind == kTileKnotKind) {
        if (live) {
            const uint32_t lc = item.y & 0xffffu, ld = item.y >> 16;
            float4 p0 = lds[la], p1 = lds[lb], p2 = lds[lc], p3 = lds[ld];
            if (projectKnot(p0, p1, p2, p3, knotCC[c], lam, dt, failure)) {
                lds[la] = p0;
                lds[lb] = p1;
                lds[lc] = p2;
                lds[ld] = p3;
            }
        }
        if (tileItemOwner(kind, item)) out[c] = lam;
    } else if (kind == kTileDistanceKind) {
        if (live) {
            float4 p = lds[la], q = lds[lb];
            const float2 rc = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
            if (projectDistance(p, q, rc, lam, invDt2, failure)) {
                lds[la] = p;
                lds[lb] = q;
            }
        }
        if (tileItemOwner(kind, item)) out[c] = lam;
    } else if (kind == kTileBendKind) {
        if (live) {
            float4 p = lds[la], q = lds[lb];
            const float2 cc = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
            if (projectBend(p, q, cc, lam, invDt2, ground, floorZ, failure)) {
                lds[la] = p;
                lds[lb] = q;
            }
        }
        if (tileItemOwner(kind, item)) out[c] = lam;
    }
    return false;
}

// The register slots of a distance or bend pass, two at a time where any
// lane has both: both slots' knots are read, both projections computed
// (branch-free, so the compiler interleaves the two dependency chains) and
// both written back. A pass is one colour, whose constraints share no knot,
// so reading the second slot's knots before the first slot's are written
// changes nothing. Slots past the pass hold zero items (knots 0 and 0).
__device__ __forceinline__ void tilePairSlots(float4* lds, bool bend, const uint4* items, float* lam, uint32_t j0,
                                              uint32_t last, uint32_t stride, float* out, uint32_t* failure,
                                              bool live, float invDt2, bool ground, float floorZ) {
    uint32_t fail = 0u;
    auto project = [&](float4& p, float4& q, const uint4& item, float& l, uint32_t& f) {
        const float2 c = make_float2(__uint_as_float(item.z), __uint_as_float(item.w));
        return bend ? projectBendCore(p, q, c, l, invDt2, ground, floorZ, f)
                    : projectDistanceCore(p, q, c, l, invDt2, f);
    };
    auto finish = [&](bool valid, bool applied, const uint4& item, const float4& p, const float4& q, float l,
                      uint32_t f, uint32_t m) {
        const bool run = live && valid;
        fail |= run ? f : 0u;
        if (run && applied) {
            lds[item.x & 0xffffu] = p;
            lds[item.x >> 16] = q;
            lam[m] = l;
        }
        if (valid && (item.y >> 31) != 0u) out[item.y & 0x7fffffffu] = lam[m];
    };
#pragma unroll
    for (uint32_t m0 = 0; m0 < kTilePrefetch; m0 += 2) {
        const bool v0 = j0 + m0 * stride < last, v1 = j0 + (m0 + 1) * stride < last;
        if (!__any(v0)) break;
        const uint4 i0 = items[m0];
        float4 p0 = lds[i0.x & 0xffffu], q0 = lds[i0.x >> 16];
        float l0 = lam[m0];
        uint32_t f0 = 0u;
        if (__any(v1)) {
            const uint4 i1 = items[m0 + 1];
            float4 p1 = lds[i1.x & 0xffffu], q1 = lds[i1.x >> 16];
            float l1 = lam[m0 + 1];
            uint32_t f1 = 0u;
            const bool a0 = project(p0, q0, i0, l0, f0);
            const bool a1 = project(p1, q1, i1, l1, f1);
            finish(v0, a0, i0, p0, q0, l0, f0, m0);
            finish(v1, a1, i1, p1, q1, l1, f1, m0 + 1);
        } else {
            const bool a0 = project(p0, q0, i0, l0, f0);
            finish(v0, a0, i0, p0, q0, l0, f0, m0);
        }
    }
    if (fail != 0u) recordFailure(failure, fail);
}

// The first kTilePrefetch item slots of a thread in every pass are software
// pipelined: the items of pass k + 2 and the multipliers of pass k + 1 are
// requested before pass k runs, so their latency overlaps the LDS work and
// the barrier instead of following each other. Slots beyond those (tiles
// with more than kTilePrefetch x workgroup items in one pass) load in place.
//
// Staged = true: the tile's items of every pass, and the input multiplier of
// each, are copied to LDS before the first pass, all requests in flight at
// once (items first, then the multipliers they index, alongside the knot
// gather), so the passes wait on no global load at all. Used when the
// tile's whole program fits in LDS and the autotuner finds it faster: tiles
// with one workgroup per CU, whose passes would otherwise each wait out a
// global load latency.
//
// Strain programs stop after the first sweep (all distance colours) that
// moves no knot of the tile: strain passes keep no state besides the knots,
// so every later sweep would repeat the same checks on the same positions
// and move nothing either. The owned knots are the same bytes.
// Strain = true: a strain-limit program (strain passes only, early exit);
// false: an iteration program (distance, knot and bend passes).
template <bool Strain, bool Staged>
__global__ void k_tile_program(TileProgramArgs a) {
    extern __shared__ float4 lds[];
    __shared__ SphereWindow window;
    __shared__ float4 scratch[2 * kMaxBlockWaves];
    __shared__ uint32_t starts[kMaxTilePasses + 1];
    __shared__ uint32_t kinds[kMaxTilePasses];
    // Multiplier buffers by pass kind (distance, knot, bend), read from LDS:
    // choosing among fields of the kernel argument by a run-time kind would
    // copy the whole argument block to scratch.
    __shared__ const float* lambdaIn[3];
    __shared__ float* lambdaOut[3];
    static_assert(sizeof(window) + sizeof(scratch) + sizeof(starts) + sizeof(kinds) + sizeof(lambdaIn) +
                      sizeof(lambdaOut) <= kTileStaticLds,
                  "static LDS beyond the host's reservation");
    if (threadIdx.x == 0) {
        lambdaIn[kTileDistanceKind] = a.distanceLambdaIn;
        lambdaIn[kTileKnotKind] = a.knotLambdaIn;
        lambdaIn[kTileBendKind] = a.bendLambdaIn;
        lambdaOut[kTileDistanceKind] = a.distanceLambdaOut;
        lambdaOut[kTileKnotKind] = a.knotLambdaOut;
        lambdaOut[kTileBendKind] = a.bendLambdaOut;
    }
    const float2* __restrict__ knotCC = a.knotCosineCompliance;
    uint32_t* const failure = a.failure;
    // Workgroups are dealt to the eight XCDs in turn (workgroup b runs on XCD
    // b mod 8); each XCD takes a run of consecutive tiles instead, so that
    // most of a tile's halo knots were written by tiles on its own XCD and
    // come from its L2. Any bijection gives the same bytes.
    const uint32_t xcd = blockIdx.x % kXcds, turn = blockIdx.x / kXcds;
    const uint32_t tile = xcd * (gridDim.x / kXcds) + min(xcd, gridDim.x % kXcds) + turn;
    if constexpr (Strain) {
        if (a.prevChanged != nullptr) {
            const uint32_t n0 = a.tileNbrStart[tile], n1 = a.tileNbrStart[tile + 1];
            bool moved = false;
            for (uint32_t k = n0 + threadIdx.x; k < n1; k += blockDim.x) moved |= a.prevChanged[a.tileNbrs[k]] != 0u;
            if (__syncthreads_or(moved ? 1 : 0) == 0) {
                if (threadIdx.x == 0) {
                    a.tileChanged[tile] = 0u;
                    atomicAdd(a.roundStats, 1ull);
                }
                return;
            }
        }
    }
    const uint32_t base = a.loadStart[tile];
    const uint32_t loaded = a.loadStart[tile + 1] - base;
    const uint32_t owned = a.ownedCount[tile];
    const uint32_t passes = a.passes;
    float4* sphereCache = lds + loaded;
    float4* culled = sphereCache + a.sphereCount;
    const bool collides = a.leadingCollide != 0 || a.trailingCollides != 0;
    const uint32_t* __restrict__ itemStart = a.itemStart + tile * passes;
    const uint4* __restrict__ items = a.items;
    // The pass kinds go to LDS: indexing the kernel argument's array with a
    // loop counter would copy the whole argument block to scratch.
    for (uint32_t k = threadIdx.x; k < passes; k += blockDim.x) kinds[k] = a.kinds[k];
    for (uint32_t k = threadIdx.x; k <= passes; k += blockDim.x) starts[k] = itemStart[k];
    // Scalars every pass reads, requested before the gather.
    const bool live = *a.ok != 0u;
    const float dt = a.cfg->gravityAndTimestep.w;
    const float invDt2 = 1.0f / (dt * dt);
    const bool ground = a.cfg->constraintCounts.z != 0u;
    const float floorZ = a.cfg->clothMaterial.x;
    // Items of passes 0 and 1 and multipliers of pass 0, in flight together
    // with the knot gather.
    const uint32_t t = threadIdx.x, stride = blockDim.x;
    uint4 cur[kTilePrefetch], next[kTilePrefetch], ahead[kTilePrefetch];
    float lam[kTilePrefetch], nextLam[kTilePrefetch];
    uint4* const stagedItems = reinterpret_cast<uint4*>(lds + a.stageOffset);
    float* const stagedLambdas = reinterpret_cast<float*>(stagedItems + a.stageItems);
    const uint32_t itemBase = itemStart[0];
    // Ids of the thread's first loaded knots (the first kept for the write-back).
    uint32_t ids[kTileKnotSlots];
    // The register slots' knots as read (strain programs compare the owned
    // knots with them when tracking changes).
    float4 orig[kTileKnotSlots];
    if constexpr (Staged) {
        // Two phases, each with every request of the workgroup in flight at
        // once: (1) the knot ids into registers and the items of every pass
        // into LDS; (2) after a barrier, the knot positions by id (with the
        // touched-mask words of the self-contact corrections) and the input
        // multiplier of every item. Slot arrays are filled for every slot
        // (indices clamped to the last element) so they stay in registers;
        // only live slots are stored.
        const uint32_t count = itemStart[passes] - itemBase;
        constexpr uint32_t kItemSlots = 8;
        tileLoadKnotIds(a, base, loaded, t, stride, ids);
        for (uint32_t j0 = t; j0 < count; j0 += kItemSlots * stride) {
            uint4 v[kItemSlots];
#pragma unroll
            for (uint32_t u = 0; u < kItemSlots; ++u)
                v[u] = tileItemAt<Strain>(items, itemBase + min(j0 + u * stride, count - 1u));
#pragma unroll
            for (uint32_t u = 0; u < kItemSlots; ++u)
                if (j0 + u * stride < count) stagedItems[j0 + u * stride] = v[u];
        }
        if (collides)
            for (uint32_t k = t; k < a.sphereCount; k += stride) sphereCache[k] = a.spheres[k];
        __syncthreads();
        tileGatherKnots(lds, a, base, loaded, owned, t, stride, live, ids, orig, [&]() {
            if constexpr (!Strain) {
                for (uint32_t j0 = t; j0 < count; j0 += kItemSlots * stride) {
                    float v[kItemSlots];
#pragma unroll
                    for (uint32_t u = 0; u < kItemSlots; ++u) {
                        const uint32_t j = min(j0 + u * stride, count - 1u);
                        uint32_t k = 0;
                        while (starts[k + 1] - itemBase <= j) ++k;
                        const uint32_t kind = kinds[k];
                        v[u] = lambdaIn[kind][tileItemConstraint(kind, stagedItems[j])];
                    }
#pragma unroll
                    for (uint32_t u = 0; u < kItemSlots; ++u)
                        if (j0 + u * stride < count) stagedLambdas[j0 + u * stride] = v[u];
                }
            }
        });
        __syncthreads();
    } else {
        // Knot ids first: the gather waits on them, while the first items
        // arrive meanwhile; the multipliers they index are requested with
        // the knot positions.
        tileLoadKnotIds(a, base, loaded, t, stride, ids);
        tileLoadItems<Strain>(items, itemStart, passes, 0, t, stride, cur);
        tileLoadItems<Strain>(items, itemStart, passes, 1, t, stride, next);
        tileGatherKnots(lds, a, base, loaded, owned, t, stride, live, ids, orig, [&]() {
            if constexpr (!Strain) {
                const uint32_t kind0 = a.kinds[0];
                const float* in0 = kind0 == kTileDistanceKind ? a.distanceLambdaIn
                                                              : (kind0 == kTileKnotKind ? a.knotLambdaIn : a.bendLambdaIn);
                tileLoadLambdas(in0, itemStart, passes, 0, kind0, t, stride, cur, lam);
            }
            if (collides)
                for (uint32_t k = t; k < a.sphereCount; k += stride) sphereCache[k] = a.spheres[k];
        });
        __syncthreads();
    }
    bool windowReady = false;
    if (live && a.leadingCollide != 0)
        tileCollide(lds, tile, loaded, owned, a, ground, floorZ, sphereCache, culled, window, scratch, windowReady);
    bool sweepMoved = false;
    for (uint32_t k = 0; k < passes; ++k) {
        const uint32_t kind = kinds[k];
        float* out = Strain ? nullptr : lambdaOut[kind];
        const uint32_t first = starts[k], last = starts[k + 1];
        bool moved = false;
        if constexpr (Staged) {
            for (uint32_t j = first - itemBase + t; j < last - itemBase; j += stride) {
                const float lj = Strain ? 0.0f : stagedLambdas[j];
                moved |= tileItem<Strain>(lds, kind, stagedItems[j], lj, out, knotCC, failure, live, dt, invDt2,
                                          ground, floorZ);
            }
        } else {
            if constexpr (!Strain)
                if (k + 1 < passes)
                    tileLoadLambdas(lambdaIn[kinds[k + 1]], itemStart, passes, k + 1, kinds[k + 1], t, stride, next,
                                    nextLam);
            tileLoadItems<Strain>(items, itemStart, passes, k + 2, t, stride, ahead);
            const float* in = Strain ? nullptr : lambdaIn[kind];
            if (!Strain && (kind == kTileDistanceKind || kind == kTileBendKind)) {
                tilePairSlots(lds, kind == kTileBendKind, cur, lam, first + t, last, stride, out, failure, live,
                              invDt2, ground, floorZ);
            } else {
#pragma unroll
                for (uint32_t m = 0; m < kTilePrefetch; ++m)
                    if (first + t + m * stride < last)
                        moved |= tileItem<Strain>(lds, kind, cur[m], lam[m], out, knotCC, failure, live, dt,
                                                  invDt2, ground, floorZ);
            }
            for (uint32_t j = first + t + kTilePrefetch * stride; j < last; j += stride) {
                const uint4 item = tileItemAt<Strain>(items, j);
                const float lj = Strain ? 0.0f : in[tileItemConstraint(kind, item)];
                moved |= tileItem<Strain>(lds, kind, item, lj, out, knotCC, failure, live, dt, invDt2, ground,
                                          floorZ);
            }
        }
        if constexpr (Strain) {
            sweepMoved |= __syncthreads_or(moved ? 1 : 0) != 0;
            if ((k + 1) % a.sweepPasses == 0) {
                if (!sweepMoved) break;
                sweepMoved = false;
            }
        } else {
            (void)moved;
            __syncthreads();
        }
        if constexpr (!Staged) {
#pragma unroll
            for (uint32_t m = 0; m < kTilePrefetch; ++m) {
                cur[m] = next[m];
                next[m] = ahead[m];
                lam[m] = nextLam[m];
            }
        }
    }
    if (live)
        for (uint32_t pass = 0; pass < a.trailingCollides; ++pass)
            tileCollide(lds, tile, loaded, owned, a, ground, floorZ, sphereCache, culled, window, scratch,
                        windowReady);
    if constexpr (Strain) {
        if (a.tileChanged != nullptr) {
            // Owned knots as written against as read, bit for bit (the
            // register slots' reads, or the input again past them).
            bool differs = false;
#pragma unroll
            for (uint32_t u = 0; u < kTileKnotSlots; ++u) {
                const uint32_t i = t + u * stride;
                if (i < owned) {
                    const float4 now = lds[i];
                    differs |= __float_as_uint(now.x) != __float_as_uint(orig[u].x) ||
                               __float_as_uint(now.y) != __float_as_uint(orig[u].y) ||
                               __float_as_uint(now.z) != __float_as_uint(orig[u].z) ||
                               __float_as_uint(now.w) != __float_as_uint(orig[u].w);
                }
            }
            for (uint32_t i = t + kTileKnotSlots * stride; i < owned; i += stride) {
                const float4 now = lds[i], was = a.posIn[a.loadIds[base + i]];
                differs |= __float_as_uint(now.x) != __float_as_uint(was.x) ||
                           __float_as_uint(now.y) != __float_as_uint(was.y) ||
                           __float_as_uint(now.z) != __float_as_uint(was.z) ||
                           __float_as_uint(now.w) != __float_as_uint(was.w);
            }
            const bool changed = __syncthreads_or(differs ? 1 : 0) != 0;
            if (t == 0) {
                a.tileChanged[tile] = changed ? 1u : 0u;
                atomicAdd(a.roundStats + 1, 1ull);
            }
        }
    }
    // The first slot's id is still in a register (the others would cost the
    // passes registers): most tiles own fewer knots than the workgroup has
    // threads.
    if (t < owned) a.posOut[ids[0]] = lds[t];
    for (uint32_t i = t + stride; i < owned; i += stride) a.posOut[a.loadIds[base + i]] = lds[i];
}

// The substep's sphere window of every tile of a program: the box of the
// tile's loaded knots at the substep's start and the spheres that can act
// inside it grown by the pad (buildSphereWindow). One workgroup of 1024
// threads per tile; dynamic LDS: the spheres and the culled list
// (2 * sphereCount float4).
__global__ void k_tile_windows(const NumiClothBagGPUConfig* __restrict__ cfg, const float4* __restrict__ pos,
                               const uint32_t* __restrict__ loadStart, const uint32_t* __restrict__ loadIds,
                               const float4* __restrict__ spheres, uint32_t sphereCount,
                               SphereWindow* __restrict__ windows, float4* __restrict__ windowSpheres) {
    extern __shared__ float4 sphereCache[];
    __shared__ SphereWindow window;
    __shared__ float4 scratch[2 * kMaxBlockWaves];
    float4* culled = sphereCache + sphereCount;
    const uint32_t tile = blockIdx.x;
    for (uint32_t k = threadIdx.x; k < sphereCount; k += blockDim.x) sphereCache[k] = spheres[k];
    V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
    // Eight knots per thread and round, every id then every position in
    // flight together (the loads, not the reduction, bound this kernel).
    constexpr uint32_t kSlots = 8;
    const uint32_t begin = loadStart[tile], end = loadStart[tile + 1];
    for (uint32_t i0 = begin + threadIdx.x; i0 < end; i0 += kSlots * blockDim.x) {
        uint32_t ids[kSlots];
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) ids[u] = loadIds[min(i0 + u * blockDim.x, end - 1u)];
        float4 q[kSlots];
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) q[u] = pos[ids[u]];
#pragma unroll
        for (uint32_t u = 0; u < kSlots; ++u) {
            const float4 p = q[u];
            if (!isfinite(p.z)) continue;  // slots past the end repeat the last knot
            lo = v3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
            hi = v3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
        }
    }
    __syncthreads();
    buildSphereWindow(lo, hi, cfg->clothMaterial.x, sphereCache, sphereCount, culled, window, scratch);
    if (threadIdx.x == 0) windows[tile] = window;
    for (uint32_t k = threadIdx.x; k < window.count; k += blockDim.x)
        windowSpheres[static_cast<size_t>(tile) * sphereCount + k] = culled[k];
}

__global__ void k_copy_float4(const float4* __restrict__ from, float4* __restrict__ to, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) to[i] = from[i];
}

__global__ void k_copy_float(const float* __restrict__ from, float* __restrict__ to, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) to[i] = from[i];
}

}  // namespace numi_dev
