#pragma once

// Fruit/yarn round-robin solves (reference solve_yarn_batch,
// src/metal/ClothBag.metal:1675-1786, and apply_yarn_friction, :3024-3158).
//
// The reference order: for each distance colour batch of n segments, for
// phase p = 0..n-1, fruit f visits segment first + (p + f) mod n; within a
// phase the fruits touch distinct segments of one colour (disjoint knots),
// so the phase is parallel and the phases are serial. Fruits f and f + n
// would meet the same segment in one phase; they take turns (rounds), lower
// fruit first.
//
// Almost every visit is a no-op (the fruit is nowhere near the segment), and
// a no-op changes nothing, so the same result comes from running only the
// visits that can act, in the same order:
//   1. candidates: every (fruit, segment) visit that can act, counted per
//      phase and scattered into phase order (count / scan / scatter in LDS);
//   2. wave 0 runs the candidate visits phase by phase (one lane per fruit),
//      knots in LDS, records staged in LDS;
//   3. commit.
// For the swept mode and for friction a visit acts exactly when its record
// says so (impact flag, stored impulse), so the candidate set is exact. For
// the present mode a visit acts when the fruit overlaps the segment at that
// moment; candidates are the visits within `margin` of contact at the start
// of the call, and after the run every other visit is checked to have stayed
// out of contact: its start distance minus the distance its fruit and its
// knots have moved in this call (sums of correction lengths) must still be
// >= contact. If any visit fails that check, or the candidates overflow the
// LDS staging, the call is run again visit by visit from the unchanged
// global state. Either way the result is the reference's sequential one.

#include "kernels_bag.hpp"
#include "kernels_self.hpp"

namespace numi_dev {

constexpr uint32_t kYarnSlots = 1024;   // staged candidate visits per call
constexpr uint32_t kYarnThreads = 1024;
// Present-mode kept list: pairs within kYarnListMargins candidate margins of
// contact at the build, at most kYarnListCapacity of them (more: no list).
constexpr float kYarnListMargins = 6.0f;
constexpr uint32_t kYarnListCapacity = 8192;

// Orders LDS accesses of one wavefront (lanes of wave 0 exchange knots
// through LDS between phases).
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fields of a fruit/yarn record a visit updates.
struct YarnStage {
    mr_float4 normalImpulse;  // fruitNormalAndImpulse
    mr_float4 segment;        // segmentImpulse
    uint32_t responses;     // control.w
    uint32_t acted;
};

struct YarnArgs {
    const NumiClothBagGPUConfig* cfg;
    const uint32_t* ok;
    uint32_t mode;                       // 1 swept impacts, 0 present overlap
    const uint4* batches;                // (first, count) per distance colour
    uint32_t batchCount;
    const uint2* segmentBatch;           // per segment: (first, count) of its colour
    float4* pos;
    float4* vel;                         // friction only
    uint32_t particles;
    const uint2* ij;
    uint32_t segments;
    NumiClothBagGPUFruit* fruits;
    uint32_t fruitCount;
    NumiClothBagGPUYarnContact* contacts;
    NumiClothBagGPUFrictionStatus* status;  // friction only
    float margin;                        // present mode candidate margin
    uint32_t* stats;                     // [reruns visit by visit, most candidates, list builds, kept-list
                                         // solves], or null
    uint32_t* failure;
    // The grips and fruit-pair colours the schedule runs just before (0:
    // none): the solve runs them first, in its own workgroup.
    NumiClothBagGPUGrip* grips;
    uint32_t gripCount;
    const uint2* pairColours;
    const uint32_t* pairColourOf;        // per fruit pair: its colour
    uint32_t pairColourCount;
    NumiClothBagGPUFruitPair* fruitPairs;
    uint32_t fruitPairCount;
    // Present-mode kept list (null: none): header (pairs, valid), the listed
    // pair indices, the knots and fruits (centre, limit) of the build.
    uint32_t* listHeader;
    uint4* listPairs;                    // (pair index, segment knot a, knot b, -)
    float4* listKnots;
    float4* listFruits;
    uint32_t listCapacity;
    float listMargin;                    // listed: within this of contact at the build
    uint32_t visitByVisit;               // 1: run every call visit by visit (the reference's literal order)
    // The present self-contact solve that follows in the schedule (null
    // grid: none), run by this workgroup on the committed positions.
    const SelfGrid* selfGrid;
    const uint4* selfPairKnots;
    const float4* selfPrev;
    FixedAccumulator* selfAcc;
    uint32_t* selfTouched;
    NumiClothBagGPUSelfImpulse* selfImpulses;
    NumiClothBagGPUSelfStatus* selfStatus;
};

// The self-contact solve a fruit/yarn call carries, after its positions are
// written (every thread of the workgroup).
__device__ __forceinline__ void yarnThenSelf(const YarnArgs& a) {
    if (a.selfGrid == nullptr) return;
    __syncthreads();
    selfSolve(a.cfg, 0u, a.selfGrid, a.selfPairKnots, a.pos, a.selfPrev, a.selfAcc, a.selfTouched, a.selfImpulses,
              a.selfStatus, 0u, 1u);
}

__device__ __forceinline__ uint32_t phaseBit(uint32_t f, uint32_t s, uint2 batch) {
    const uint32_t j = s - batch.x;
    const uint32_t fm = f % batch.y;
    return batch.x + (j + batch.y - fm) % batch.y;
}

// The substep constants a visit reads, loaded once per call.
struct YarnVisitConstants {
    bool ground;
    float floorZ;  // the yarn radius: the knots' floor and the contact offset
    float dt;
};

__device__ __forceinline__ YarnVisitConstants yarnVisitConstants(const YarnArgs& a) {
    return YarnVisitConstants{a.cfg->constraintCounts.z != 0u, a.cfg->clothMaterial.x, a.cfg->gravityAndTimestep.w};
}

// The visit (fruit, segment) of the solve: updates fruit, knots a and b (in
// LDS) and the staged record; returns the correction lengths it applied.
__device__ __forceinline__ float yarnVisit(const YarnArgs& a, const YarnVisitConstants& k,
                                           NumiClothBagGPUFruit& fruit, const NumiClothBagGPUYarnContact& record,
                                           YarnStage& stage, float4* knots, uint2 e, float* moved, bool& acted) {
    acted = false;
    const bool ground = k.ground;
    const float floorZ = k.floorZ;
    const float dt = k.dt;
    V3 n = v3(0.0f, 0.0f, 0.0f);
    float wa = 0.0f, wb = 0.0f, correction = 0.0f;
    if (a.mode == 1u) {
        if (record.control.y != 0u) {
            n = xyz(record.sweptNormalAndAdvance);
            wb = record.weightsAndTime.y;
            wa = 1.0f - wb;
            correction = record.sweptNormalAndAdvance.w;
        }
    } else {
        const V3 ka = xyz(knots[e.x]), kb = xyz(knots[e.y]);
        const V3 centre = xyz(fruit.positionAndInverseMass);
        const PointSample s = samplePointSegment(centre, ka, kb);
        const float target = fruit.previousAndRadius.w + floorZ;
        if (s.distance < target) {
            n = s.distance < 1.0e-12f ? degenerateYarnNormal(ka, kb) : (s.closest - centre) / s.distance;
            wb = s.weight;
            wa = 1.0f - wb;
            correction = target - s.distance;
        }
    }
    if (!(correction > 0.0f)) return 0.0f;
    bool wrote = false;
    float4 ka = knots[e.x], kb = knots[e.y];
    const V3 c0 = xyz(fruit.positionAndInverseMass);
    const float lambda = yarnCorrection(ka, kb, fruit, wa, wb, n, correction, ground, floorZ, a.failure, wrote);
    float travel = length(xyz(fruit.positionAndInverseMass) - c0);
    if (wrote) {
        moved[e.x] += length(xyz(ka) - xyz(knots[e.x]));
        moved[e.y] += length(xyz(kb) - xyz(knots[e.y]));
        knots[e.x] = ka;
        knots[e.y] = kb;
    }
    if (lambda > 0.0f) {
        const float impulse = lambda / dt;
        setXyz(stage.normalImpulse, xyz(stage.normalImpulse) - n * impulse);
        stage.normalImpulse.w += impulse;
        stage.segment.x += wa * impulse;
        stage.segment.y += wb * impulse;
        if (a.mode == 1u) {
            stage.segment.z = record.weightsAndTime.z;
            stage.segment.w += correction;
        }
        stage.responses += 1u;
        acted = true;
    }
    return travel;
}

// Whether point c is certainly farther than `limit` from the segment of
// midpoint m and half length h: |c - m| - h > limit plus a slack that covers
// the rounding of the distances (point-segment distances computed in float
// are within a few ulps of the scene's coordinates of the true ones). The
// exact test (samplePointSegment) is then sure to fail, so the candidate
// listing and the check of the unlisted visits skip it; both ways give the
// same candidates. Non-finite input is never "beyond".
__device__ __forceinline__ bool beyondSegment(V3 c, V3 m, float h, float limit) {
    const V3 q = c - m;
    const float scale = fmaxf(fmaxf(fmaxf(fabsf(c.x), fabsf(c.y)), fmaxf(fabsf(c.z), fabsf(m.x))),
                              fmaxf(fabsf(m.y), fabsf(m.z)));
    const float r = limit + h + 1.0e-4f + 1.0e-5f * scale;
    return dot(q, q) > r * r;
}

// Exclusive scan of counts[0..n) in place by the whole workgroup; returns
// the total (same value in every thread).
__device__ uint32_t blockExclusiveScan(uint32_t* counts, uint32_t n, uint32_t* scratch) {
    const uint32_t tid = threadIdx.x, threads = blockDim.x;
    const uint32_t chunk = (n + threads - 1) / threads;
    const uint32_t begin = min(n, tid * chunk), end = min(n, begin + chunk);
    uint32_t sum = 0;
    for (uint32_t i = begin; i < end; ++i) sum += counts[i];
    scratch[tid] = sum;
    __syncthreads();
    for (uint32_t offset = 1; offset < threads; offset <<= 1) {
        const uint32_t add = tid >= offset ? scratch[tid - offset] : 0u;
        __syncthreads();
        scratch[tid] += add;
        __syncthreads();
    }
    uint32_t running = scratch[tid] - sum;
    const uint32_t total = scratch[threads - 1];
    for (uint32_t i = begin; i < end; ++i) {
        const uint32_t c = counts[i];
        counts[i] = running;
        running += c;
    }
    __syncthreads();
    return total;
}

// The visit-by-visit schedule on wave 0 (the fallback, and the reference's
// literal order): knots in LDS, records read and written in global memory.
__device__ void yarnDense(const YarnArgs& a, float4* knots, float* moved) {
    const uint32_t lane = threadIdx.x;
    const bool live = lane < a.fruitCount;
    const YarnVisitConstants constants = yarnVisitConstants(a);
    NumiClothBagGPUFruit fruit{};
    if (live) fruit = a.fruits[lane];
    for (uint32_t batch = 0; batch < a.batchCount; ++batch) {
        const uint4 b = a.batches[batch];
        const uint32_t rounds = (a.fruitCount + b.y - 1) / b.y;
        for (uint32_t step = 0; step < b.y * rounds; ++step) {
            const uint32_t phase = step / rounds, round = step % rounds;
            if (live && lane / b.y == round) {
                const uint32_t segment = b.x + (phase + lane) % b.y;
                const uint32_t index = lane * a.segments + segment;
                NumiClothBagGPUYarnContact c = a.contacts[index];
                YarnStage stage{c.fruitNormalAndImpulse, c.segmentImpulse, c.control.w, 0u};
                bool acted = false;
                yarnVisit(a, constants, fruit, c, stage, knots, a.ij[segment], moved, acted);
                if (acted) {
                    c.fruitNormalAndImpulse = stage.normalImpulse;
                    c.segmentImpulse = stage.segment;
                    c.control.w = stage.responses;
                    a.contacts[index] = c;
                }
            }
            waveSync();
        }
    }
    if (live) {
        if (!finite3(xyz(fruit.positionAndInverseMass))) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        a.fruits[lane] = fruit;
    }
}

// Largest value over the workgroup of a per-thread non-negative float (NaN
// as +inf): wave maxima into perWave[], which the caller reads after its
// next barrier.
__device__ __forceinline__ void waveMaxInto(float v, float* perWave) {
    v = v < INFINITY ? v : INFINITY;
    for (int offset = 32; offset > 0; offset >>= 1) v = fmaxf(v, __shfl_xor(v, offset));
    if ((threadIdx.x & 63u) == 0u) perWave[threadIdx.x / 64u] = v;
}

__device__ __forceinline__ float maxAbs3(V3 v) {
    const float m = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.z));
    return m < INFINITY ? m : INFINITY;
}

// One solve_yarn_batch call over all distance colours. One workgroup of
// kYarnThreads; dynamic LDS: knots[P], moved[P], visitors[D], visited[D]
// (per segment: fruits with a candidate visit on it, fruits whose visit has
// run).
//
// Present mode with a kept list (listPairs non-null): the list holds the
// (fruit, segment) visits that were within `listMargin` of contact when it
// was built, with the knots and the fruit centres of then and, per fruit,
// the limit it was listed at. The distance of a point to a segment changes
// by no more than the point and the segment's endpoints move, so a visit
// off the list is still beyond contact while its fruit's room (limit minus
// the current contact distance) exceeds the largest knot drift plus the
// largest fruit drift since the build, plus a rounding slack. A solve whose
// start drift leaves half the room lists its candidates from the kept list
// alone; otherwise it tests every visit, as without a list, and rebuilds the
// list from its start state. The check after the run covers the visits off
// the list with the same bound (now including the knots' and fruits' travel
// in the call), and falls back to checking every visit when it fails.
__global__ void __launch_bounds__(kYarnThreads) k_yarn_solve(YarnArgs a) {
    extern __shared__ float4 yarnLds[];
    __shared__ unsigned long long rankKeys[kYarnSlots];  // staging: (colour, fruit, phase) per candidate
    __shared__ uint32_t candidates[kYarnSlots];     // pair indices, unordered
    __shared__ uint32_t slotPair[kYarnSlots];        // pair indices by colour, fruit, phase
    __shared__ uint32_t slotFruit[kYarnSlots];       // the fruit of each slot
    __shared__ uint32_t slotColourEnd[kYarnSlots];   // end of the slots of the slot's colour
    __shared__ uint32_t firstOwn[32];                // walk: each fruit's first slot in the colour
    __shared__ uint32_t slotEarlier[kYarnSlots];     // fruits visiting the slot's segment before it
    __shared__ uint2 slotEnds[kYarnSlots];           // the slot's segment knots
    __shared__ YarnStage slotStage[kYarnSlots];
    __shared__ NumiClothBagGPUFruit fruitStart[32];
    __shared__ float4 keptFruit[32];                 // kept list: centre, limit per fruit at the build
    __shared__ float fruitMoved[32];
    __shared__ mr_float4 fruitStaged[32];
    __shared__ float waveDrift[kYarnThreads / 64];   // kept list: knot drift since the build
    __shared__ float waveScale[kYarnThreads / 64];   // largest knot coordinate magnitude
    __shared__ float waveMoved[kYarnThreads / 64];   // largest knot travel in the call
    __shared__ float fruitBounds[4];  // kept list: fruit drift, kept room, build room, fruit scale
    __shared__ uint32_t fruitBoundFlags;  // bit 0: build room finite, bit 1: kept room finite
    __shared__ uint32_t candidateCount;
    __shared__ uint32_t listCount;  // pairs a rebuild listed
    __shared__ uint32_t keptCount;  // pairs on the kept list, ~0u when there is none
    __shared__ uint32_t verdict;    // 0 sparse result stands, 1 rerun visit by visit
    if (*a.ok == 0u) return;
    const uint32_t tid = threadIdx.x, threads = blockDim.x;
    const uint32_t P = a.particles, D = a.segments, F = a.fruitCount;
    const bool listed = a.mode == 0u && a.listPairs != nullptr;
    // The kept list's entry of this thread, read ahead (within capacity).
    const uint4 prefetched = listed && tid < a.listCapacity ? a.listPairs[tid] : make_uint4(0u, 0u, 0u, 0u);
    if (a.gripCount != 0u) {
        // solve_grip: one grip per thread, each on its own knot.
        for (uint32_t i = tid; i < a.gripCount; i += threads)
            projectGrip(a.cfg, a.pos, a.grips, i, a.gripCount, P, a.failure);
        __syncthreads();
    }
    if (a.pairColourCount != 0u)
        fruitPairColours(a.cfg, a.pairColours, a.pairColourOf, a.pairColourCount, a.fruits, F, a.fruitPairs,
                         a.fruitPairCount, a.failure, fruitStart);
    float4* knots = yarnLds;
    float* moved = reinterpret_cast<float*>(knots + P);
    uint32_t* visitors = reinterpret_cast<uint32_t*>(moved + P);
    uint32_t* visited = visitors + D;
    float drift = 0.0f, scale = 0.0f;
    for (uint32_t i = tid; i < P; i += threads) {
        const float4 x = a.pos[i];
        knots[i] = x;
        moved[i] = 0.0f;
        if (listed) {
            const float d = length(xyz(x) - xyz(a.listKnots[i]));
            drift = fmaxf(drift, d < INFINITY ? d : INFINITY);
            scale = fmaxf(scale, maxAbs3(xyz(x)));
        }
    }
    if (listed) {
        waveMaxInto(drift, waveDrift);
        waveMaxInto(scale, waveScale);
    }
    for (uint32_t i = tid; i < D; i += threads) {
        visitors[i] = 0u;
        visited[i] = 0u;
    }
    const float yarnRadius = a.cfg->clothMaterial.x;
    if (tid < 64) {
        // Wave 0, one lane per fruit (the pair colours left the fruits in
        // fruitStart); kept list: the fruits' drift since the build, the
        // room per fruit (its listing limit minus its contact distance) of
        // the kept list and of a rebuild, their coordinate magnitude.
        float fruitDrift = 0.0f, keptRoom = INFINITY, buildRoom = INFINITY, fruitScale = 0.0f;
        bool buildable = true, keepable = true;
        if (tid < F) {
            const NumiClothBagGPUFruit fr = a.pairColourCount != 0u ? fruitStart[tid] : a.fruits[tid];
            fruitStart[tid] = fr;
            fruitMoved[tid] = 0.0f;
            if (listed) {
                const float4 kept = a.listFruits[tid];
                keptFruit[tid] = kept;
                const V3 c = xyz(fr.positionAndInverseMass);
                const float target = fr.previousAndRadius.w + yarnRadius;
                buildRoom = (target + a.listMargin) - target;
                keptRoom = kept.w - target;
                buildable = buildRoom == buildRoom;
                keepable = keptRoom == keptRoom;
                fruitScale = maxAbs3(c);
                const float d = length(c - xyz(kept));
                fruitDrift = d < INFINITY ? d : INFINITY;
            }
        }
        if (listed) {
            for (int offset = 32; offset > 0; offset >>= 1) {
                fruitDrift = fmaxf(fruitDrift, __shfl_xor(fruitDrift, offset));
                fruitScale = fmaxf(fruitScale, __shfl_xor(fruitScale, offset));
                keptRoom = fminf(keptRoom, __shfl_xor(keptRoom, offset));
                buildRoom = fminf(buildRoom, __shfl_xor(buildRoom, offset));
            }
            const bool allBuildable = __all(buildable), allKeepable = __all(keepable);
            if (tid == 0) {
                fruitBounds[0] = fruitDrift;
                fruitBounds[1] = keptRoom;
                fruitBounds[2] = buildRoom;
                fruitBounds[3] = fruitScale;
                fruitBoundFlags = (allBuildable ? 1u : 0u) | (allKeepable ? 2u : 0u);
            }
        }
    }
    if (tid == 0) {
        candidateCount = 0u;
        listCount = 0u;
        keptCount = listed && a.listHeader[1] != 0u ? a.listHeader[0] : ~0u;
        verdict = 0u;
    }
    __syncthreads();
    // Kept list: 0 none, 1 candidates from the kept list, 2 rebuild. Every
    // thread reaches the same decision from the same LDS values.
    uint32_t listState = 0u;
    float keptDrift = 0.0f, room = INFINITY, slack = 0.0f;
    if (listed) {
        float knotDrift = 0.0f;
        for (uint32_t w = 0; w < threads / 64u; ++w) {
            knotDrift = fmaxf(knotDrift, waveDrift[w]);
            scale = fmaxf(scale, waveScale[w]);
        }
        const float fruitDrift = fruitBounds[0], keptRoom = fruitBounds[1];
        room = fruitBounds[2];
        scale = fmaxf(scale, fruitBounds[3]);
        const bool buildable = (fruitBoundFlags & 1u) != 0u;
        const bool keepable = (fruitBoundFlags & 2u) != 0u && keptCount != ~0u;
        slack = 1.0e-4f + 1.0e-5f * scale;
        listState = buildable ? 2u : 0u;
        if (buildable && keepable) {
            keptDrift = knotDrift + fruitDrift;
            // Off-list visits start beyond the candidate margin, and half
            // the room is left for the travel of this call.
            if (keptDrift + 2.0f * slack + a.margin <= 0.5f * keptRoom) {
                listState = 1u;
                room = keptRoom;
            }
        }
    }
    // 1. candidate visits: counted per phase and listed (start state in LDS).
    // A visit can act in the swept mode when its record's impact flag is
    // set, in the present mode when the fruit is within margin of contact
    // at the start of the call. Nothing depends on the order of the listing.
    auto stageCandidate = [&](uint32_t f, uint32_t s, uint32_t idx) {
        atomicOr(&visitors[s], 1u << f);
        const uint32_t k = atomicAdd(&candidateCount, 1u);
        if (k < kYarnSlots) candidates[k] = idx;
    };
    // Kept list: thread t keeps entry t and its segment's start knots in
    // registers for the check after the run.
    uint4 heldEntry = make_uint4(~0u, 0u, 0u, 0u);
    V3 heldK0 = v3(0.0f, 0.0f, 0.0f), heldK1 = heldK0;
    if (listState == 1u) {
        for (uint32_t k = tid; k < keptCount; k += threads) {
            const uint4 entry = k == tid ? prefetched : a.listPairs[k];
            const uint32_t idx = entry.x;
            const uint32_t f = idx / D, s = idx - f * D;
            const V3 k0 = xyz(knots[entry.y]), k1 = xyz(knots[entry.z]);
            if (k == tid) {
                heldEntry = entry;
                heldK0 = k0;
                heldK1 = k1;
            }
            const NumiClothBagGPUFruit& fr = fruitStart[f];
            const float limit = fr.previousAndRadius.w + yarnRadius + a.margin;
            if (samplePointSegment(xyz(fr.positionAndInverseMass), k0, k1).distance < limit)
                stageCandidate(f, s, idx);
        }
    } else {
        // A thread takes segments and tests every fruit against each, so a
        // segment's colour and endpoints are read once.
        for (uint32_t s = tid; s < D; s += threads) {
            const uint2 batch = a.segmentBatch[s];
            if (batch.y == 0u) continue;
            const uint2 e = a.ij[s];
            const V3 k0 = xyz(knots[e.x]), k1 = xyz(knots[e.y]);
            const V3 mid = (k0 + k1) * 0.5f;
            const float half = 0.5f * length(k1 - k0);
            for (uint32_t f = 0; f < F; ++f) {
                const uint32_t idx = f * D + s;
                bool candidate;
                if (a.mode == 1u) {
                    candidate = a.contacts[idx].control.y != 0u;
                } else {
                    const NumiClothBagGPUFruit& fr = fruitStart[f];
                    const V3 c = xyz(fr.positionAndInverseMass);
                    const float limit = fr.previousAndRadius.w + yarnRadius + a.margin;
                    if (listState == 2u) {
                        const float target = fr.previousAndRadius.w + yarnRadius;
                        const float buildLimit = target + a.listMargin;
                        if (beyondSegment(c, mid, half, buildLimit)) continue;
                        const float distance = samplePointSegment(c, k0, k1).distance;
                        if (!(distance < buildLimit)) continue;
                        const uint32_t k = atomicAdd(&listCount, 1u);
                        if (k < a.listCapacity) a.listPairs[k] = make_uint4(idx, e.x, e.y, 0u);
                        candidate = distance < limit;
                    } else {
                        candidate =
                            !beyondSegment(c, mid, half, limit) && samplePointSegment(c, k0, k1).distance < limit;
                    }
                }
                if (candidate) stageCandidate(f, s, idx);
            }
        }
        if (listState == 2u) {
            for (uint32_t i = tid; i < P; i += threads) a.listKnots[i] = knots[i];
            if (tid < F) {
                const NumiClothBagGPUFruit& fr = fruitStart[tid];
                const float target = fr.previousAndRadius.w + yarnRadius;
                a.listFruits[tid] = withXyz(make_float4(0.0f, 0.0f, 0.0f, target + a.listMargin),
                                            xyz(fr.positionAndInverseMass));
            }
        }
    }
    __syncthreads();
    // Staging: slots ordered by colour, then fruit, then phase (a fruit has
    // one visit per phase, and colours are consecutive phase ranges in
    // segment order), each with its fruit, segment knots, record, the end of
    // its colour's slots and the fruits whose visit of its segment comes
    // first in the reference order (an earlier phase, or the same phase and a
    // lower fruit: fruits f and f + n of a colour of n segments). Candidate t
    // is ranked by thread t, which counts the candidates ordered before it.
    const uint32_t total = candidateCount;
    if (total > kYarnSlots || a.visitByVisit != 0u) {
        if (tid == 0) verdict = 1u;
    } else {
        // key: the colour's first segment, the fruit, the phase within the
        // colour (a colour's first segment is below 2^27 and its phases below
        // 2^32 - 1 for any scene the LDS holds).
        uint32_t idx = 0u, f = 0u, s = 0u, phase = 0u;
        uint2 batch = make_uint2(0u, 0u);
        unsigned long long key = ~0ull;
        if (tid < total) {
            idx = candidates[tid];
            f = idx / D;
            s = idx - f * D;
            batch = a.segmentBatch[s];
            phase = phaseBit(f, s, batch);
            key = (static_cast<unsigned long long>(batch.x) << 37) | (static_cast<unsigned long long>(f) << 32) |
                  (phase - batch.x);
            rankKeys[tid] = key;
        }
        __syncthreads();
        if (tid < total) {
            const unsigned long long colourLimit = (static_cast<unsigned long long>(batch.x) + 1ull) << 37;
            uint32_t rank = 0u, colourEnd = 0u;
#pragma unroll 8
            for (uint32_t j = 0; j < total; ++j) {
                const unsigned long long other = rankKeys[j];
                rank += other < key ? 1u : 0u;
                colourEnd += other < colourLimit ? 1u : 0u;
            }
            slotColourEnd[rank] = colourEnd;
            slotPair[rank] = idx;
            slotFruit[rank] = f;
            slotEnds[rank] = a.ij[s];
            uint32_t earlier = 0u;
            for (uint32_t others = visitors[s] & ~(1u << f); others != 0u; others &= others - 1u) {
                const uint32_t g = static_cast<uint32_t>(__ffs(static_cast<int>(others))) - 1u;
                const uint32_t pg = phaseBit(g, s, batch);
                if (pg < phase || (pg == phase && g < f)) earlier |= 1u << g;
            }
            slotEarlier[rank] = earlier;
            const NumiClothBagGPUYarnContact& c = a.contacts[idx];
            slotStage[rank] = YarnStage{c.fruitNormalAndImpulse, c.segmentImpulse, c.control.w, 0u};
        }
    }
    __syncthreads();
    // 2. candidate visits on wave 0, one lane per fruit, colour by colour
    // (a colour's slots are contiguous, each fruit's within them contiguous
    // and in phase order). A visit reads and writes only its fruit, the two
    // knots of its segment and its own staged record, and the segments of
    // one colour share no knot; so within a colour each lane runs its own
    // visits in phase order and a visit waits only for the visits of its
    // segment that the reference runs before it (slotEarlier, against
    // visited[s]). The earliest pending visit is always ready, and the
    // result is the reference's sequential one. Everything a round reads is
    // in LDS.
    if (verdict == 0u && tid < 64) {
        const uint32_t lane = tid;
        const bool live = lane < F;
        const YarnVisitConstants constants = yarnVisitConstants(a);
        NumiClothBagGPUFruit fruit{};
        if (live) fruit = fruitStart[lane];
        float travel = 0.0f;
        for (uint32_t colourStart = 0; colourStart < total;) {
            const uint32_t colourEnd = slotColourEnd[colourStart];
            // Each fruit's first slot in the colour: the slot that starts
            // its run.
            if (lane < 32) firstOwn[lane] = ~0u;
            waveSync();
            for (uint32_t j = colourStart + lane; j < colourEnd; j += 64) {
                const uint32_t g = slotFruit[j];
                if (j == colourStart || slotFruit[j - 1] != g) firstOwn[g] = j;
            }
            waveSync();
            uint32_t cursor = live ? firstOwn[lane] : ~0u;
            bool finished = cursor >= colourEnd;
            while (true) {
                if (__all(finished)) break;
                if (!finished) {
                    const uint32_t k = cursor;
                    const uint32_t idx = slotPair[k];
                    const uint32_t s = idx - lane * D;
                    if ((slotEarlier[k] & ~visited[s]) == 0u) {
                        YarnStage stage = slotStage[k];
                        bool acted = false;
                        travel += yarnVisit(a, constants, fruit, a.contacts[idx], stage, knots, slotEnds[k], moved,
                                            acted);
                        stage.acted = acted ? 1u : 0u;
                        slotStage[k] = stage;
                        visited[s] |= 1u << lane;
                        ++cursor;
                        finished = cursor >= colourEnd || slotFruit[cursor] != lane;
                    }
                }
                waveSync();
            }
            colourStart = colourEnd;
        }
        if (live) {
            fruitMoved[lane] = travel;
            fruitStaged[lane] = fruit.positionAndInverseMass;  // committed after the check below
        }
    }
    __syncthreads();
    // 3. present mode: every visit that was not run must have stayed a no-op.
    // Non-candidates that were tested started at least `margin` beyond
    // contact; if no fruit plus no knot travelled that far in this call,
    // none can have reached contact. Visits off the kept list are covered by
    // the drift bound. Otherwise check each such visit's own start distance.
    if (verdict == 0u && a.mode == 0u) {
        float local = 0.0f;
        for (uint32_t i = tid; i < P; i += threads) local = fmaxf(local, moved[i]);
        waveMaxInto(local, waveMoved);
        __syncthreads();
        float fruitMax = 0.0f, knotMax = 0.0f;
        for (uint32_t f = 0; f < F; ++f) fruitMax = fmaxf(fruitMax, fruitMoved[f] < INFINITY ? fruitMoved[f] : INFINITY);
        for (uint32_t w = 0; w < threads / 64u; ++w) knotMax = fmaxf(knotMax, waveMoved[w]);
        const bool cheap = fruitMax + knotMax + slack < a.margin;
        const bool haveList = listState == 1u || (listState == 2u && listCount <= a.listCapacity);
        const float startDrift = listState == 1u ? keptDrift : 0.0f;
        const bool offListHold = haveList && startDrift + knotMax + fruitMax + slack < room;
        // One visit that did not run: false when it may have reached contact.
        auto stayedOut = [&](uint32_t f, uint32_t s, V3 k0, V3 k1, V3 mid, float half, float knotReach) {
            if ((visitors[s] >> f) & 1u) return true;  // ran as a candidate
            const NumiClothBagGPUFruit& fr = fruitStart[f];
            const V3 c = xyz(fr.positionAndInverseMass);
            const float target = fr.previousAndRadius.w + yarnRadius;
            const float reach = fruitMoved[f] + knotReach;
            // Far beyond contact plus reach: stayed out of contact, without
            // the exact distance.
            if (beyondSegment(c, mid, half, target + reach)) return true;
            return samplePointSegment(c, k0, k1).distance - reach >= target;
        };
        bool bad = false;
        if (offListHold) {
            if (!cheap) {
                const uint32_t n = listState == 1u ? keptCount : listCount;
                for (uint32_t k = tid; k < n; k += threads) {
                    const bool held = listState == 1u && k == tid;
                    const uint4 entry = held ? heldEntry : a.listPairs[k];
                    const uint32_t f = entry.x / D, s = entry.x - f * D;
                    const V3 k0 = held ? heldK0 : xyz(a.pos[entry.y]), k1 = held ? heldK1 : xyz(a.pos[entry.z]);
                    if (!stayedOut(f, s, k0, k1, (k0 + k1) * 0.5f, 0.5f * length(k1 - k0),
                                   fmaxf(moved[entry.y], moved[entry.z])))
                        bad = true;
                }
            }
        } else if (!(cheap && listState != 1u)) {
            // Segment-major like the listing: a segment's colour, endpoints
            // and start positions are read once for all fruits.
            for (uint32_t s = tid; s < D; s += threads) {
                if (a.segmentBatch[s].y == 0u) continue;
                const uint2 e = a.ij[s];
                const V3 k0 = xyz(a.pos[e.x]), k1 = xyz(a.pos[e.y]);
                const V3 mid = (k0 + k1) * 0.5f;
                const float half = 0.5f * length(k1 - k0);
                const float knotReach = fmaxf(moved[e.x], moved[e.y]);
                for (uint32_t f = 0; f < F; ++f)
                    if (!stayedOut(f, s, k0, k1, mid, half, knotReach)) bad = true;
            }
        }
        if (bad) verdict = 1u;
    }
    __syncthreads();
    if (tid == 0) {
        if (a.stats != nullptr) {
            atomicAdd(&a.stats[0], verdict);
            atomicMax(&a.stats[1], total);
            atomicAdd(&a.stats[2], listState == 2u ? 1u : 0u);
            atomicAdd(&a.stats[3], listState == 1u ? 1u : 0u);
        }
        if (listState == 2u) {
            a.listHeader[0] = min(listCount, a.listCapacity);
            a.listHeader[1] = listCount <= a.listCapacity ? 1u : 0u;
        }
    }
    if (verdict == 0u) {
        for (uint32_t i = tid; i < P; i += threads) a.pos[i] = knots[i];
        for (uint32_t k = tid; k < total; k += threads) {
            const YarnStage st = slotStage[k];
            if (st.acted == 0u) continue;
            NumiClothBagGPUYarnContact& c = a.contacts[slotPair[k]];
            c.fruitNormalAndImpulse = st.normalImpulse;
            c.segmentImpulse = st.segment;
            c.control.w = st.responses;
        }
        if (tid < F) {
            NumiClothBagGPUFruit fr = fruitStart[tid];
            fr.positionAndInverseMass = fruitStaged[tid];
            if (!finite3(xyz(fr.positionAndInverseMass))) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
            a.fruits[tid] = fr;
        }
        yarnThenSelf(a);
        return;
    }
    // Rerun visit by visit from the unchanged global state.
    for (uint32_t i = tid; i < P; i += threads) {
        knots[i] = a.pos[i];
        moved[i] = 0.0f;
    }
    __syncthreads();
    if (tid < 64) yarnDense(a, knots, moved);
    __syncthreads();
    for (uint32_t i = tid; i < P; i += threads) a.pos[i] = knots[i];
    yarnThenSelf(a);
}

// Friction of one fruit/yarn record (reference apply_yarn_friction,
// src/metal/ClothBag.metal:3024-3158); velocities in LDS.
__device__ __forceinline__ void yarnFrictionVisit(const YarnArgs& a, NumiClothBagGPUFruit& fruit,
                                                  const NumiClothBagGPUYarnContact& c, float4* velocities, uint2 e) {
    const float mu = a.cfg->fruitMaterial.w;
    const bool ground = a.cfg->constraintCounts.z != 0u;
    const float jn = c.fruitNormalAndImpulse.w;
    const float nl = length(xyz(c.fruitNormalAndImpulse));
    if (!(jn > 0.0f) || !(nl > 1.0e-10f)) return;
    const V3 n = xyz(c.fruitNormalAndImpulse) / nl;
    float wx = c.segmentImpulse.x / jn, wy = c.segmentImpulse.y / jn;
    const float ws = wx + wy;
    if (!(ws > 1.0e-12f)) return;
    wx = wx / ws;
    wy = wy / ws;
    float4 va = velocities[e.x], vb = velocities[e.y];
    const float4 xa = a.pos[e.x], xb = a.pos[e.y];
    const V3 yarnV = xyz(va) * wx + xyz(vb) * wy;
    const V3 offset = n * -fruit.previousAndRadius.w;
    const V3 ballV = xyz(fruit.velocityAndGroundImpulse) + cross(xyz(fruit.angularVelocity), offset);
    const V3 rel = ballV - yarnV;
    const V3 tv = rel - n * dot(rel, n);
    const float slip = length(tv);
    if (!(slip > 1.0e-10f)) return;
    const V3 tangent = tv / slip;
    const V3 lever = cross(offset, tangent);
    float denom = fruit.positionAndInverseMass.w + fruitInverseInertia(fruit) * dot(lever, lever);
    V3 ra = tangent * xa.w;
    V3 rb = tangent * xb.w;
    if (ground) {
        if (xa.z <= a.cfg->clothMaterial.x + 1.0e-6f && ra.z < 0.0f) ra.z = 0.0f;
        if (xb.z <= a.cfg->clothMaterial.x + 1.0e-6f && rb.z < 0.0f) rb.z = 0.0f;
    }
    denom += dot(tangent, ra) * wx * wx + dot(tangent, rb) * wy * wy;
    if (!(denom > 0.0f) || !isfinite(denom)) return;
    const float limit = mu * jn;
    const float jt = fminf(slip / denom, limit);
    if (!(jt > 0.0f)) return;
    applyFruitImpulse(fruit, tangent * -jt, offset);
    velocities[e.x] = withXyz(va, xyz(va) + ra * (jt * wx));
    velocities[e.y] = withXyz(vb, xyz(vb) + rb * (jt * wy));
    recordFriction(a.status, 1, jt, limit);
}

// Yarn friction over all colours: a visit acts only if its record carries a
// normal impulse, so exactly those visits run, in phase order.
__global__ void __launch_bounds__(kYarnThreads) k_yarn_friction_solve(YarnArgs a) {
    extern __shared__ float4 yarnLds[];
    __shared__ uint32_t scanScratch[kYarnThreads];
    __shared__ uint32_t slotPair[kYarnSlots];
    __shared__ uint32_t slotPhase[kYarnSlots];
    __shared__ uint32_t overflow;
    if (*a.ok == 0u) return;
    const float mu = a.cfg->fruitMaterial.w;
    if (!(mu >= 0.0f) || !isfinite(mu)) {
        if (threadIdx.x == 0) recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
        return;
    }
    const uint32_t tid = threadIdx.x, threads = blockDim.x;
    const uint32_t P = a.particles, D = a.segments, F = a.fruitCount;
    float4* velocities = yarnLds;
    uint32_t* counts = reinterpret_cast<uint32_t*>(velocities + P);
    for (uint32_t i = tid; i < P; i += threads) velocities[i] = a.vel[i];
    for (uint32_t i = tid; i < D; i += threads) counts[i] = 0u;
    if (tid == 0) overflow = 0u;
    __syncthreads();
    const uint32_t pairs = F * D;
    auto acts = [&](uint32_t idx) {
        return a.segmentBatch[idx % D].y != 0u && a.contacts[idx].fruitNormalAndImpulse.w > 0.0f;
    };
    for (uint32_t idx = tid; idx < pairs; idx += threads) {
        const uint32_t f = idx / D, s = idx % D;
        if (acts(idx)) atomicAdd(&counts[phaseBit(f, s, a.segmentBatch[s])], 1u);
    }
    __syncthreads();
    const uint32_t total = blockExclusiveScan(counts, D, scanScratch);
    if (total > kYarnSlots) {
        if (tid == 0) overflow = 1u;
    } else {
        for (uint32_t idx = tid; idx < pairs; idx += threads) {
            const uint32_t f = idx / D, s = idx % D;
            if (!acts(idx)) continue;
            const uint32_t phase = phaseBit(f, s, a.segmentBatch[s]);
            const uint32_t slot = atomicAdd(&counts[phase], 1u);
            slotPair[slot] = idx;
            slotPhase[slot] = phase;
        }
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t lane = tid;
        const bool live = lane < F;
        NumiClothBagGPUFruit fruit{};
        if (live) fruit = a.fruits[lane];
        if (overflow != 0u) {
            // every visit, in the reference order
            for (uint32_t batch = 0; batch < a.batchCount; ++batch) {
                const uint4 b = a.batches[batch];
                const uint32_t rounds = (F + b.y - 1) / b.y;
                for (uint32_t step = 0; step < b.y * rounds; ++step) {
                    const uint32_t phase = step / rounds, round = step % rounds;
                    if (live && lane / b.y == round) {
                        const uint32_t segment = b.x + (phase + lane) % b.y;
                        yarnFrictionVisit(a, fruit, a.contacts[lane * D + segment], velocities, a.ij[segment]);
                    }
                    waveSync();
                }
            }
        } else {
            // the non-empty phase groups only, in phase order
            for (uint32_t first = 0; first < total;) {
                const uint32_t phase = slotPhase[first];
                const uint32_t last = counts[phase];
                const uint32_t n = a.segmentBatch[phase].y;
                const uint32_t rounds = (F + n - 1) / n;
                for (uint32_t round = 0; round < rounds; ++round) {
                    if (live && lane / n == round) {
                        for (uint32_t k = first; k < last; ++k) {
                            const uint32_t idx = slotPair[k];
                            if (idx / D != lane) continue;
                            yarnFrictionVisit(a, fruit, a.contacts[idx], velocities, a.ij[idx % D]);
                        }
                    }
                    waveSync();
                }
                first = last;
            }
        }
        if (live) {
            if (!finite4(fruit.velocityAndGroundImpulse) || !finite4(fruit.angularVelocity)) {
                recordFailure(a.failure, NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE);
            } else {
                a.fruits[lane] = fruit;
            }
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < P; i += threads) a.vel[i] = velocities[i];
}

}  // namespace numi_dev
