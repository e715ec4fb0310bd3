// Host runtime and C ABI of the MI355X cloth substep loop.
//
// create(): host validation of the ABI v13 arrays, AoS -> SoA conversion,
//           device-resident buffers (reference runGPU allocates and copies
//           per call, tools/cloth_bag_metal.mm:4019-4177; here state stays
//           on the device between calls).
// step():   uploads the per-substep configs, resets the device substep
//           cursor and replays one captured hipGraph per substep. The graph
//           holds the reference's per-substep dispatch order
//           (tools/cloth_bag_metal.mm:4211-4709) for the families the scene
//           has; empty families are left out of the graph.
// read():   SoA -> AoS readback (reference :4720-4757).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels_core.hpp"
#include "kernels_diag.hpp"
#include "kernels_self.hpp"
#include "kernels_tile.hpp"
#include "kernels_bag.hpp"
#include "kernels_yarn.hpp"
#include "kernels_adjoint.hpp"
#include "kernels_point.hpp"
#include "kernels_certificate.hpp"
#include "tile_plan.hpp"
#include "numi_cloth_hip.h"
#include "numi_cloth_scene.h"
#include "scene_internal.hpp"

using namespace numi_hip;
using namespace numi_dev;

namespace {

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        const hipError_t status_ = (expr);                                                    \
        if (status_ != hipSuccess) {                                                          \
            fail(NUMI_CLOTH_HIP_ERROR_DEVICE, std::string(#expr " failed: ") +                \
                                                  hipGetErrorString(status_));                \
        }                                                                                     \
    } while (0)

// Fixed number of reconciliation / final / certificate passes of the
// reference schedule (tools/cloth_bag_metal.mm:47-49).
constexpr uint32_t kReconciliationPasses = 8;
constexpr uint32_t kFinalContactPasses = 2;
constexpr uint32_t kCertificatePasses = 8;

// Substeps of the first step() call that time graph replays against direct
// launches: one replay that uploads the graph, then two of each.
constexpr uint32_t kLaunchProbeSubsteps = 5;

template <typename T>
struct DeviceArray {
    T* ptr = nullptr;
    std::size_t count = 0;
    void allocate(std::size_t n) {
        release();
        count = n;
        if (n == 0) return;
        HIP_OK(hipMalloc(&ptr, n * sizeof(T)));
    }
    // Allocates when the array holds fewer than n elements (contents are not
    // kept); a larger array is reused as is.
    void ensure(std::size_t n) {
        if (count < n) allocate(n);
    }
    void upload(const T* host, std::size_t n, hipStream_t stream) {
        if (n == 0) return;
        HIP_OK(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, stream));
    }
    void download(T* host, std::size_t n, hipStream_t stream) const {
        if (n == 0) return;
        HIP_OK(hipMemcpyAsync(host, ptr, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    }
    void downloadRange(T* host, std::size_t first, std::size_t n, hipStream_t stream) const {
        if (n == 0) return;
        HIP_OK(hipMemcpyAsync(host, ptr + first, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    }
    void release() {
        if (ptr != nullptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    ~DeviceArray() { release(); }
};

struct Range {
    uint32_t first;
    uint32_t count;
};

uint32_t blocksFor(uint32_t n) { return std::max(1u, (n + kBlock - 1) / kBlock); }

}  // namespace

struct NumiClothHipSolver {
    hipStream_t stream = nullptr;
    int device = 0;

    // Counts and topology (host copies keep the static record fields).
    NumiClothBagGPUConfig initialConfig{};
    uint32_t P = 0, D = 0, G = 0, K = 0, B = 0, S = 0;
    std::vector<NumiClothBagGPUDistance> hostDistances;
    std::vector<NumiClothBagGPUBend> hostBends;
    std::vector<NumiClothBagGPUKnot> hostKnots;
    std::vector<Range> distanceColours, bendColours, knotColours;
    std::vector<uint32_t> cuff;
    NumiClothHipSceneExtension extension{};
    bool driven = false;

    // Device state.
    DeviceArray<float4> pos, prev, vel, anchor, motion;
    DeviceArray<uint2> dIJ, bEnds;
    DeviceArray<float2> dRestCompliance, bChordCompliance, kCosineCompliance;
    DeviceArray<float> dLambda, dMaxLength, bLambda, kLambda;
    DeviceArray<uint4> kQuad;
    DeviceArray<NumiClothBagGPUGrip> grips;
    DeviceArray<uint32_t> dCuff;
    DeviceArray<float4> spheres;
    DeviceArray<NumiClothBagGPUConfig> configs;
    DeviceArray<double> times;      // per substep: time at the substep end
    DeviceArray<SubstepSlot> slot;
    DeviceArray<uint32_t> failure;
    DeviceArray<uint32_t> roundFlags;  // per-tile change flags of the reconciliation and certificate rounds' strain programs
    DeviceArray<unsigned long long> roundStats;  // reconciliation tile-rounds since create: skipped, run
    DeviceArray<NumiClothBagGPUSelfStatus> selfStatus;
    DeviceArray<NumiClothBagGPUFrictionStatus> frictionStatus;
    DeviceArray<NumiClothBagGPUAerodynamicsStatus> aeroStatus;
    DeviceArray<NumiClothBagGPUReleaseStatus> releaseStatus;
    DeviceArray<float4> diagPartials, diagResult;
    // Certificate loop (extension.certificatePasses != 0): per-workgroup
    // residual maxima, the reduced residuals and their pinned host copy.
    DeviceArray<float4> certificatePartials, certificateResult;
    float4* certificateHost = nullptr;
    uint32_t certificateBlocks = 0;
    uint32_t certificateLast = 0;  // passes of the last step() call

    // Produce-bag families (fruits, fruit pairs, fruit/yarn records, air).
    uint32_t F = 0, FP = 0, Y = 0;
    bool yarnAir = false;
    std::vector<Range> pairColours;
    DeviceArray<uint2> pairColourRanges;
    DeviceArray<uint32_t> pairColourOf;  // per fruit pair: the index of its colour
    DeviceArray<uint2> segmentBatch;  // per distance: (first, count) of its colour
    DeviceArray<uint32_t> yarnStats;  // sparse yarn solve: reruns, most candidates, list builds, kept-list solves
    // Present-mode fruit/yarn kept list: header (pairs, valid), pairs, the
    // knots and fruits of its build; off with NUMI_CLOTH_HIP_YARN_LIST=0.
    bool yarnList = true;
    bool yarnVisitByVisit = false;  // NUMI_CLOTH_HIP_YARN_VISIT_BY_VISIT=1: every solve in the reference's order
    // Bag-sized scenes run each iteration's present self-contact solve in the
    // fruit/yarn solve's workgroup (NUMI_CLOTH_HIP_SELF_IN_YARN=0: its own
    // launch); selfSolvedInYarn: the last fruit/yarn launch carried it.
    bool selfInYarn = false;
    bool selfSolvedInYarn = false;
    DeviceArray<uint32_t> yarnListHeader;
    DeviceArray<uint4> yarnListPairs;
    DeviceArray<float4> yarnListKnots, yarnListFruits;
    std::vector<uint32_t> rimHost;
    DeviceArray<NumiClothBagGPUFruit> fruits;
    DeviceArray<NumiClothBagGPUFruitPair> fruitPairs;
    DeviceArray<NumiClothBagGPUYarnContact> yarnContacts;
    DeviceArray<uint4> distanceBatchesDevice;
    DeviceArray<float4> aeroForces, aeroReduction;
    DeviceArray<uint32_t> rim, mouthClearance;

    // Hash self-contact (present when extension.selfContact != 0).
    bool self = false;
    uint32_t entryCapacity = 0, pairCapacity = 0, entryTiles = 0, segmentBlocks = 0;
    uint32_t cellCapacity = 0;          // per-cell counters of the counting sort (0: radix sort only)
    DeviceArray<uint32_t> cellSlots;
    DeviceArray<uint32_t> localStart, localList;
    DeviceArray<uint32_t> localFixed;   // two-hop lists in localWidth slots per knot (0: not used)
    uint32_t localWidth = 0;
    DeviceArray<float4> boxLo, boxHi, partialLo, partialHi;
    DeviceArray<SelfGrid> grid;
    DeviceArray<uint32_t> segmentCounts, entryCounts, scanSums, totals;
    DeviceArray<uint32_t> keysA, valsA, keysB, valsB, histogram;
    DeviceArray<uint2> pairs;
    DeviceArray<uint4> pairKnots;  // per candidate: the endpoint knots of both segments
    DeviceArray<uint4> foundPairs;  // run walks: (entry, rank, s, t) of every candidate, in any order
    // Two fixed-point accumulators and touched masks (one bit per knot with a
    // pending correction), used by alternate self-contact solves: a tile
    // program that applies one buffer's corrections cannot clear it (other
    // tiles read it in the same launch), so the next solve uses the other.
    DeviceArray<FixedAccumulator> accumulator[2];
    DeviceArray<uint32_t> touched[2];
    uint32_t selfBuffer = 0;   // buffer of the next self-contact solve (record time)
    int pendingSelfApply = -1; // buffer whose corrections the next launch applies
    int dirtySelf = -1;        // buffer applied by tile programs, still to clear
    DeviceArray<NumiClothBagGPUSelfImpulse> selfImpulses;  // per-candidate impulse aggregate of the substep
    // Point self-collision (present when extension.pointSelfCollision != 0):
    // hashed knot grid (bucket counts -> ends, knot ids by bucket), its own
    // fixed-point accumulators, the contact count of the last pass.
    bool pointSelf = false;
    uint32_t pointBuckets = 0;
    DeviceArray<uint32_t> pointEnds, pointIds, pointTouched, pointContacts, pointBucketCount, pointScanSums,
        pointTotal;
    DeviceArray<FixedAccumulator> pointAcc;
    DeviceArray<SubstepSlot> debugSlot;

    // LDS tile programs (NUMI_CLOTH_HIP_TILES=0 keeps the per-colour launches).
    struct TileProgram {
        bool valid = false;
        uint32_t tiles = 0, passes = 0, maxLoaded = 0, block = 256, owned = 0, halo = 0;
        uint32_t maxItems = 0;     // most items of one tile, all passes
        bool stageable = false;    // the whole program of every tile fits in LDS
        bool staged = false;       // run the staged kernel
        bool hasDistance = false, hasBend = false, hasKnot = false;
        bool packedStrain = false;  // 8-byte strain items (k_tile_program<true, *>)
        std::vector<uint32_t> kinds;
        DeviceArray<uint32_t> loadStart, loadIds, ownedCount, itemStart;
        DeviceArray<uint4> items;
        // Per tile, the tiles owning its loaded knots (itself included).
        DeviceArray<uint32_t> nbrStart, nbrs;
        // Per tile, the substep's sphere window and its list (windowed scenes).
        DeviceArray<SphereWindow> windows;
        DeviceArray<float4> windowSpheres;
        double algorithmicBytes = 0.0;
        uint64_t computed = 0, unique = 0;
    };
    bool useTiles = false;
    TileProgram tileEven, tileOdd, tileStrain;
    uint32_t tileOwned = 0, tileBlock = 256, tileCapacity = 0, tileStrainSweeps = ~0u;
    bool tileStaged = false;            // the autotuner's choice for every program that can stage
    uint64_t ldsBytes = 0;              // dynamic LDS available to a tile program
    std::vector<float4> initialPositions;  // create-time layout used to order tiles
    DeviceArray<float4> posAlt;
    DeviceArray<float> dLambdaAlt, bLambdaAlt, kLambdaAlt;
    // Buffers holding the live state while a substep is recorded.
    float4* posCur = nullptr;
    float* dLamCur = nullptr;
    float* bLamCur = nullptr;
    float* kLamCur = nullptr;

    // A substep is launched either as a replay of its captured hipGraph or
    // kernel by kernel on the stream; the bytes are the same. The first
    // step() of at least kLaunchProbeSubsteps substeps times both ways on the
    // live scene and keeps the faster (on an MI355X with an idle host the two
    // are within 2 % for the 1M and 256^2 sheets and the bag; a busy host
    // slows the direct launches first). NUMI_CLOTH_HIP_GRAPHS=0 / 1 fixes the
    // choice (=0 also under profilers that do not follow graph launches).
    bool useGraphs = true;
    bool launchDecided = false;
    float launchProbeGraphMs = 0.0f, launchProbeDirectMs = 0.0f;
    uint32_t sphereWindowMin = kSphereWindowMin;

    // Pinned staging for per-step uploads.
    NumiClothBagGPUConfig* stagingConfigs = nullptr;
    double* stagingTimes = nullptr;
    uint32_t stagingCapacity = 0;
    double time = 0.0;  // simulated time at the end of the last substep

    // One captured substep; rebuilt when (iterations, strainSweeps) change.
    hipGraphExec_t graph = nullptr;
    uint32_t graphIterations = ~0u;
    uint32_t graphSweeps = ~0u;

    ~NumiClothHipSolver() {
        if (graph != nullptr) (void)hipGraphExecDestroy(graph);
        if (stagingConfigs != nullptr) (void)hipHostFree(stagingConfigs);
        if (stagingTimes != nullptr) (void)hipHostFree(stagingTimes);
        if (certificateHost != nullptr) (void)hipHostFree(certificateHost);
        if (stream != nullptr) (void)hipStreamDestroy(stream);
    }

    const NumiClothBagGPUConfig* cfg() const { return &slot.ptr->config; }
    const uint32_t* ok() const { return &slot.ptr->ok; }

    // ---------------------------------------------------------- launches --

    void launchBegin(bool fuseAdvance, float dampingFactor) {
        BeginArgs a{};
        a.cfg = cfg();
        a.ok = ok();
        a.pos = posCur;
        a.prev = prev.ptr;
        a.vel = vel.ptr;
        a.anchor = driven ? anchor.ptr : nullptr;
        a.motion = motion.ptr;
        a.drive = make_float4(extension.driveAmplitude, extension.driveOmega, extension.drivePhase, 0.0f);
        a.time = &slot.ptr->time;
        a.dampingDt = initialConfig.gravityAndTimestep.w;
        a.distanceLambda = dLamCur;
        a.bendLambda = bLamCur;
        a.knotLambda = kLamCur;
        a.grips = grips.ptr;
        a.particles = P;
        a.distances = D;
        a.bends = B;
        a.knots = K;
        a.grips_ = G;
        a.dampingFactor = dampingFactor;
        a.advance = fuseAdvance ? 1u : 0u;
        a.failure = failure.ptr;
        const uint32_t n = std::max({P, D, B, K, G});
        k_begin<<<blocksFor(n), kBlock, 0, stream>>>(a);
    }

    void launchDistances() {
        flushCollide();
        for (const Range& r : distanceColours) {
            k_distance<<<blocksFor(r.count), kBlock, 0, stream>>>(cfg(), ok(), r.first, r.count, posCur, dIJ.ptr,
                                                                  dRestCompliance.ptr, dLamCur, failure.ptr);
        }
    }

    void launchKnots() {
        flushCollide();
        for (const Range& r : knotColours) {
            k_knot<<<blocksFor(r.count), kBlock, 0, stream>>>(cfg(), ok(), r.first, r.count, posCur, kQuad.ptr,
                                                              kCosineCompliance.ptr, kLamCur, failure.ptr);
        }
    }

    void launchBends() {
        flushCollide();
        for (const Range& r : bendColours) {
            k_bend<<<blocksFor(r.count), kBlock, 0, stream>>>(cfg(), ok(), r.first, r.count, posCur, bEnds.ptr,
                                                              bChordCompliance.ptr, bLamCur, failure.ptr);
        }
    }

    void launchStrain(uint32_t sweeps) {
        flushCollide();
        for (uint32_t sweep = 0; sweep < sweeps; ++sweep) {
            for (const Range& r : distanceColours) {
                k_strain<<<blocksFor(r.count), kBlock, 0, stream>>>(ok(), r.first, r.count, posCur, dIJ.ptr,
                                                                    dMaxLength.ptr, failure.ptr);
            }
        }
    }

    bool groundOrColliders() const {
        return initialConfig.constraintCounts.z != 0u || extension.staticSphereCount != 0u;
    }

    // ---- bag families (no-ops when the scene has no fruits) ----
    uint32_t fruitLanes() const { return std::max(64u, (F + 63u) / 64u * 64u); }

    void launchFruitPairs() {
        if (pairColours.empty()) return;
        flushCollide();
        k_fruit_pairs<<<1, kFruitPairThreads, 0, stream>>>(cfg(), ok(), pairColourRanges.ptr, pairColourOf.ptr,
                                                          static_cast<uint32_t>(pairColours.size()), fruits.ptr, F,
                                                          fruitPairs.ptr, FP, failure.ptr);
    }

    YarnArgs yarnArgs(uint32_t mode, float4* positions) const {
        YarnArgs a{};
        a.cfg = cfg();
        a.ok = ok();
        a.mode = mode;
        a.batches = distanceBatchesDevice.ptr;
        a.batchCount = static_cast<uint32_t>(distanceColours.size());
        a.segmentBatch = segmentBatch.ptr;
        a.pos = positions;
        a.vel = vel.ptr;
        a.particles = P;
        a.ij = dIJ.ptr;
        a.segments = D;
        a.fruits = fruits.ptr;
        a.fruitCount = F;
        a.contacts = yarnContacts.ptr;
        a.status = frictionStatus.ptr;
        // Present-mode candidates: within one yarn radius of contact.
        a.margin = initialConfig.clothMaterial.x;
        a.stats = yarnStats.ptr;
        a.failure = failure.ptr;
        a.visitByVisit = yarnVisitByVisit ? 1u : 0u;
        if (yarnList && mode == 0u) {
            a.listHeader = yarnListHeader.ptr;
            a.listPairs = yarnListPairs.ptr;
            a.listKnots = yarnListKnots.ptr;
            a.listFruits = yarnListFruits.ptr;
            a.listCapacity = static_cast<uint32_t>(yarnListPairs.count);
            a.listMargin = kYarnListMargins * a.margin;
        }
        return a;
    }

    void launchYarn(uint32_t mode, bool withPairs = false, bool withGrips = false, bool withSelf = false) {
        if (Y == 0) return;
        flushCollide();
        const size_t lds = static_cast<size_t>(P) * (sizeof(float4) + sizeof(float)) + static_cast<size_t>(D) * 8;
        YarnArgs a = yarnArgs(mode, posCur);
        if (withGrips && G != 0) {
            a.grips = grips.ptr;
            a.gripCount = G;
        }
        if (withPairs) {
            a.pairColours = pairColourRanges.ptr;
            a.pairColourOf = pairColourOf.ptr;
            a.pairColourCount = static_cast<uint32_t>(pairColours.size());
            a.fruitPairs = fruitPairs.ptr;
            a.fruitPairCount = FP;
        }
        if (withSelf && self) {
            const uint32_t b = selfBuffer;
            if (dirtySelf == static_cast<int>(b)) {
                clearSelf(b);
                dirtySelf = -1;
            }
            a.selfGrid = grid.ptr;
            a.selfPairKnots = pairKnots.ptr;
            a.selfPrev = prev.ptr;
            a.selfAcc = accumulator[b].ptr;
            a.selfTouched = touched[b].ptr;
            a.selfImpulses = selfImpulses.ptr;
            a.selfStatus = selfStatus.ptr;
            selfSolvedInYarn = true;
        }
        k_yarn_solve<<<1, kYarnThreads, lds, stream>>>(a);
    }

    // The grips (withGrips), the fruit-pair colours and the present
    // fruit/yarn solve that follows them, in one workgroup launch when the
    // scene has fruit/yarn records.
    // withSelf: the present self-contact solve that follows joins the launch
    // (launchSelf then launches only what comes after the solve).
    void launchFruitPairsAndYarn(bool withGrips = false, bool withSelf = false) {
        if (Y != 0) {
            launchYarn(0, !pairColours.empty(), withGrips, withSelf);
            return;
        }
        if (withGrips) launchGrips();
        launchFruitPairs();
        launchYarn(0);
    }

    // Whether a ground/sphere launch also runs the fruit ground pass (fruits
    // and the ground on): k_collide takes it then.
    bool collideTakesFruitGround() const { return F != 0 && initialConfig.constraintCounts.z != 0u; }

    void launchBuildYarnContacts() {
        if (Y == 0) return;
        flushCollide();
        k_yarn_contacts_build<<<blocksFor(Y), kBlock, 0, stream>>>(cfg(), ok(), posCur, prev.ptr, dIJ.ptr, D,
                                                                   fruits.ptr, F, yarnContacts.ptr, failure.ptr);
    }

    void launchFruitGround() {
        if (F == 0 || initialConfig.constraintCounts.z == 0u) return;
        flushCollide();
        k_fruit_ground<<<blocksFor(F), kBlock, 0, stream>>>(cfg(), ok(), fruits.ptr, F, failure.ptr);
    }

    // solveFinalContacts of the reference: fruit pairs, fruit/yarn, ground
    // (cloth, then fruits).
    void launchFinalContacts() {
        launchFruitPairsAndYarn();
        launchCollide(true);
        if (!collideTakesFruitGround()) launchFruitGround();
    }

    // fruitGround: the fruit ground pass that follows joins the launch (the
    // caller then skips launchFruitGround when collideTakesFruitGround()).
    void launchCollide(bool fruitGround = false) {
        flushCollide();
        collideNow(fruitGround);
    }

    // selfBufferFirst >= 0: apply (and clear) that self-contact buffer's
    // corrections in the same launch, before the pass.
    void collideNow(bool fruitGround = false, int selfBufferFirst = -1) {
        const bool withFruits = fruitGround && collideTakesFruitGround();
        if (!groundOrColliders()) {
            if (selfBufferFirst >= 0) applySelf(static_cast<uint32_t>(selfBufferFirst), posCur);
            if (withFruits) launchFruitGround();
            return;
        }
        const uint32_t n = extension.staticSphereCount;
        FixedAccumulator* acc = selfBufferFirst >= 0 ? accumulator[selfBufferFirst].ptr : nullptr;
        uint32_t* mask = selfBufferFirst >= 0 ? touched[selfBufferFirst].ptr : nullptr;
        k_collide<<<blocksFor(P), kBlock, 2 * n * sizeof(float4), stream>>>(
            cfg(), ok(), posCur, P, spheres.ptr, n, sphereWindowMin, failure.ptr, withFruits ? fruits.ptr : nullptr,
            withFruits ? F : 0u, acc, mask);
    }

    // A ground/sphere pass that follows self-contact can run at the start of
    // the next tile program instead of as a pass of its own (it acts on one
    // knot at a time, so loaded halo knots stay exact). Every other launch
    // that reads or writes positions first runs it here.
    bool pendingCollide = false;

    // Runs what was left to the next tile program: the self-contact
    // corrections, then the ground/sphere pass.
    void flushCollide() {
        if (pendingSelfApply >= 0) {
            applySelf(static_cast<uint32_t>(pendingSelfApply), posCur);
            pendingSelfApply = -1;
        }
        if (!pendingCollide) return;
        pendingCollide = false;
        collideNow();
    }

    // Exclusive scan of `capacity`-sized data whose live length is *count.
    void launchScan(uint32_t* data, const uint32_t* count, uint32_t capacity, uint32_t* total,
                    uint32_t* sums = nullptr) {
        if (sums == nullptr) sums = scanSums.ptr;
        const uint32_t tiles = std::min(kScanBlocks, std::max(1u, (capacity + kScanTile - 1) / kScanTile));
        k_scan_tiles<<<tiles, kScanThreads, 0, stream>>>(data, count, sums);
        k_scan_sums<<<1, kScanThreads, 0, stream>>>(sums, count, total);
        k_scan_add<<<tiles, kScanThreads, 0, stream>>>(data, count, sums);
    }

    // Point self-collision pass: hash the knots, solve every pair on the
    // snapshot into the point accumulators, apply the means (and clear).
    void launchPointSelf() {
        if (!pointSelf) return;
        flushCollide();
        const uint32_t mask = pointBuckets - 1u;
        k_point_clear<<<blocksFor(pointBuckets), kBlock, 0, stream>>>(pointEnds.ptr, pointBuckets, pointContacts.ptr);
        k_point_count<<<blocksFor(P), kBlock, 0, stream>>>(cfg(), ok(), posCur, P, mask, pointEnds.ptr);
        launchScan(pointEnds.ptr, pointBucketCount.ptr, pointBuckets, pointTotal.ptr, pointScanSums.ptr);
        k_point_place<<<blocksFor(P), kBlock, 0, stream>>>(cfg(), ok(), posCur, P, mask, pointEnds.ptr, pointIds.ptr);
        k_point_solve<<<blocksFor(P), kBlock, 0, stream>>>(cfg(), ok(), posCur, P, mask, pointEnds.ptr, pointIds.ptr,
                                                            localStart.ptr, localList.ptr, pointAcc.ptr,
                                                            pointTouched.ptr, pointContacts.ptr);
        const uint32_t words = (P + 31) / 32;
        k_self_apply<<<blocksFor(words), kBlock, 0, stream>>>(ok(), posCur, pointAcc.ptr, pointTouched.ptr, words,
                                                              failure.ptr);
    }

    // Broad phase: boxes, grid, cell entries, radix sort, candidate pairs.
    void launchBroadPhase(const NumiClothBagGPUConfig* config, const uint32_t* valid) {
        const float margin = extension.selfContactMargin;
        SelfGrid* g = grid.ptr;
        k_self_boxes<<<segmentBlocks, kScanThreads, 0, stream>>>(config, margin, posCur, prev.ptr, dIJ.ptr, D,
                                                                 boxLo.ptr, boxHi.ptr, partialLo.ptr,
                                                                 partialHi.ptr);
        k_self_grid<<<1, kScanThreads, 0, stream>>>(partialLo.ptr, partialHi.ptr, segmentBlocks, config, valid, D,
                                                    cellCapacity, g, failure.ptr);
        k_self_cell_counts<<<segmentBlocks, kScanThreads, 0, stream>>>(boxLo.ptr, boxHi.ptr, g, segmentCounts.ptr,
                                                                       cellSlots.ptr, failure.ptr);
        launchScan(segmentCounts.ptr, &g->segments, D, totals.ptr);
        // The entries in emission order into A; on a dense grid each also
        // draws its slot among its cell's entries (into entryCounts, free
        // until the candidate pass).
        k_self_emit_entries<<<segmentBlocks, kScanThreads, 0, stream>>>(boxLo.ptr, boxHi.ptr, g, segmentCounts.ptr,
                                                                        totals.ptr, entryCapacity, keysA.ptr,
                                                                        valsA.ptr, cellSlots.ptr, entryCounts.ptr,
                                                                        failure.ptr);
        // Grid-stride kernels over the live entries: enough blocks to fill the
        // chip, not one per capacity slot.
        const uint32_t entryBlocks = std::min<uint32_t>(8192, (entryCapacity + kScanThreads - 1) / kScanThreads);
        // Dense grid: counting sort (per-cell counts from the emit pass, one
        // scan, placement into B by cell start + slot; the ranking by segment
        // within each cell happens in the candidate pass, or in the rank pass
        // for lists searched in memory). Otherwise these exit at once and the
        // radix passes below sort A; each side's kernels exit when the other
        // sorts.
        if (cellCapacity != 0) {
            launchScan(cellSlots.ptr, &g->cellScanLength, cellCapacity, totals.ptr + 3);
            k_self_place_entries<<<entryBlocks, kScanThreads, 0, stream>>>(g, cellSlots.ptr, keysA.ptr, valsA.ptr,
                                                                           entryCounts.ptr, keysB.ptr, valsB.ptr);
            if (localWidth == 0)
                k_self_rank_entries<<<entryBlocks, kScanThreads, 0, stream>>>(g, cellSlots.ptr, keysB.ptr, valsB.ptr,
                                                                              keysA.ptr, valsA.ptr);
        }
        uint32_t* keysIn = keysA.ptr;
        uint32_t* valsIn = valsA.ptr;
        uint32_t* keysOut = keysB.ptr;
        uint32_t* valsOut = valsB.ptr;
        const uint32_t radixBlocks = std::min(kScanBlocks, entryTiles);
        for (uint32_t shift = 0; shift < 32; shift += kRadixBits) {
            k_radix_histogram<<<radixBlocks, kScanThreads, 0, stream>>>(keysIn, &g->radixCount, shift, histogram.ptr);
            launchScan(histogram.ptr, &g->histLength, kRadixBins * entryTiles, totals.ptr + 2);
            k_radix_scatter<<<radixBlocks, kScanThreads, 0, stream>>>(keysIn, valsIn, &g->radixCount, shift,
                                                                       histogram.ptr, keysOut, valsOut);
            std::swap(keysIn, keysOut);
            std::swap(valsIn, valsOut);
        }
        // Four passes: the radix sort ends in A, where the candidate pass (or
        // the rank pass) also writes the counting sort's order.
        static_assert(32 / kRadixBits % 2 == 0, "the radix sort must end in A");
        PairArgs pa{};
        pa.boxLo = boxLo.ptr;
        pa.boxHi = boxHi.ptr;
        pa.ij = dIJ.ptr;
        pa.grid = g;
        pa.keys = keysA.ptr;
        pa.vals = valsA.ptr;
        pa.placedKeys = keysB.ptr;
        pa.placedVals = valsB.ptr;
        pa.localStart = localStart.ptr;
        pa.localList = localList.ptr;
        pa.localFixed = localFixed.ptr;
        pa.counts = entryCounts.ptr;
        pa.offsets = entryCounts.ptr;
        pa.pairTotal = totals.ptr + 1;
        pa.capacity = pairCapacity;
        pa.pairs = pairs.ptr;
        pa.pairKnots = pairKnots.ptr;
        pa.impulses = selfImpulses.ptr;
        pa.failure = failure.ptr;
        pa.found = foundPairs.ptr;
        // Lists that fit 16 or 32 register slots: run walks over LDS windows
        // (one workgroup per chunk's runs) that also append the candidates,
        // then the emission pass puts them in scanned order. Longer lists:
        // one thread per entry with the lists searched in memory.
        if (localWidth != 0) {
            const uint32_t chunks = std::min<uint32_t>(4096, (entryCapacity + kRunChunk - 1) / kRunChunk);
            if (localWidth == 16) k_self_run_counts<16, kRunChunk, kRunWindow><<<chunks, kPairThreads, 0, stream>>>(pa);
            else k_self_run_counts<32, kRunChunk, kRunWindow><<<chunks, kPairThreads, 0, stream>>>(pa);
        } else {
            k_self_pair_counts<<<entryBlocks, kPairThreads, 0, stream>>>(pa);
        }
        launchScan(entryCounts.ptr, &g->sortCount, entryCapacity, totals.ptr + 1);
        if (localWidth != 0) {
            const uint32_t blocks = std::min<uint32_t>(2048, (pairCapacity + kScanThreads - 1) / kScanThreads);
            if (localWidth == 16) k_self_emit_run_pairs<16><<<blocks, kScanThreads, 0, stream>>>(pa);
            else k_self_emit_run_pairs<32><<<blocks, kScanThreads, 0, stream>>>(pa);
        } else {
            k_self_emit_pairs<<<entryBlocks, kScanThreads, 0, stream>>>(pa);
        }
    }

    // collide: run the ground/sphere pass that follows; defer: leave it to the
    // next launch (the next tile program takes it as its leading pass).
    // fruitGround: a collide run here also takes the fruit ground pass.
    void launchSelf(uint32_t mode, bool collide = false, bool defer = false, bool fruitGround = false) {
        flushCollide();
        if (!self) return;
        const uint32_t b = selfBuffer;
        if (!selfSolvedInYarn) {
            if (dirtySelf == static_cast<int>(b)) {
                clearSelf(b);
                dirtySelf = -1;
            }
            // Four workgroups per CU: 262,144 candidates per round of the
            // kernel's grid-stride loop, so a draped 1M-knot sheet (~230,000
            // candidates) waits out its chain of dependent loads (pair,
            // endpoints, positions) once, and a call with no candidate still
            // costs little more than its launch.
            const uint32_t blocks = std::min<uint32_t>(1024, std::max(1u, (pairCapacity + kBlock - 1) / kBlock));
            k_self_solve<<<blocks, kBlock, 0, stream>>>(cfg(), ok(), mode, grid.ptr, pairKnots.ptr, posCur,
                                                         prev.ptr, accumulator[b].ptr, touched[b].ptr,
                                                         selfImpulses.ptr, selfStatus.ptr);
        }
        selfSolvedInYarn = false;
        selfBuffer ^= 1u;
        const bool colliding = collide && groundOrColliders();
        if (defer) {
            pendingSelfApply = static_cast<int>(b);
            pendingCollide = colliding;
            return;
        }
        if (colliding) {
            collideNow(fruitGround, static_cast<int>(b));  // the corrections first, in the same launch
        } else {
            applySelf(b, posCur);
            if (fruitGround && collideTakesFruitGround()) launchFruitGround();
        }
    }

    void applySelf(uint32_t b, float4* target) {
        const uint32_t words = (P + 31) / 32;
        k_self_apply<<<blocksFor(words), kBlock, 0, stream>>>(ok(), target, accumulator[b].ptr, touched[b].ptr, words,
                                                              failure.ptr);
    }

    void clearSelf(uint32_t b) { applySelf(b, nullptr); }

    // Ends the substep's self-contact buffers: nothing pending, both clean.
    void settleSelf() {
        flushCollide();
        if (dirtySelf >= 0) clearSelf(static_cast<uint32_t>(dirtySelf));
        dirtySelf = -1;
        selfBuffer = 0;
    }

    // Self-contact friction after finalize (velocities) and the cloth/ground
    // friction fused into it, before the fruit/yarn frictions.
    void launchSelfFriction() {
        if (!self) return;
        const uint32_t blocks = std::min<uint32_t>(2048, std::max(1u, (pairCapacity + kBlock - 1) / kBlock));
        k_self_friction<<<blocks, kBlock, 0, stream>>>(cfg(), ok(), grid.ptr, pairs.ptr, dIJ.ptr, pos.ptr, vel.ptr,
                                                        selfImpulses.ptr, accumulator[0].ptr, touched[0].ptr,
                                                        frictionStatus.ptr);
        applySelf(0, vel.ptr);
    }

    // tileChanged / prevChanged: the strain programs' per-tile change flags
    // (see TileProgramArgs); a program with a leading pass skips no tile.
    void launchTileProgram(TileProgram& t, uint32_t trailingCollides, const NumiClothBagGPUConfig* config,
                           const uint32_t* valid, bool scratch = false, uint32_t* tileChanged = nullptr,
                           const uint32_t* prevChanged = nullptr) {
        TileProgramArgs a{};
        a.cfg = config;
        a.ok = valid;
        a.posIn = posCur;
        a.posOut = posCur == pos.ptr ? posAlt.ptr : pos.ptr;
        a.loadStart = t.loadStart.ptr;
        a.loadIds = t.loadIds.ptr;
        a.ownedCount = t.ownedCount.ptr;
        a.itemStart = t.itemStart.ptr;
        a.items = t.items.ptr;
        a.distanceRestCompliance = dRestCompliance.ptr;
        a.distanceMaxLength = dMaxLength.ptr;
        a.distanceLambdaIn = dLamCur;
        a.distanceLambdaOut = dLamCur == dLambda.ptr ? dLambdaAlt.ptr : dLambda.ptr;
        a.bendChordCompliance = bChordCompliance.ptr;
        a.bendLambdaIn = bLamCur;
        a.bendLambdaOut = bLamCur == bLambda.ptr ? bLambdaAlt.ptr : bLambda.ptr;
        a.knotCosineCompliance = kCosineCompliance.ptr;
        a.knotLambdaIn = kLamCur;
        a.knotLambdaOut = kLamCur == kLambda.ptr ? kLambdaAlt.ptr : kLambda.ptr;
        a.spheres = spheres.ptr;
        a.sphereCount = extension.staticSphereCount;
        a.sphereWindowMin = sphereWindowMin;
        a.passes = t.passes;
        a.sweepPasses = &t == &tileStrain ? static_cast<uint32_t>(distanceColours.size()) : 0u;
        a.leadingCollide = !scratch && pendingCollide ? 1u : 0u;
        if (!scratch) {
            pendingCollide = false;
            if (pendingSelfApply >= 0) {
                a.selfAcc = accumulator[pendingSelfApply].ptr;
                a.selfTouched = touched[pendingSelfApply].ptr;
            }
            if (dirtySelf >= 0) {
                a.clearAcc = accumulator[dirtySelf].ptr;
                a.clearTouched = touched[dirtySelf].ptr;
            }
            dirtySelf = pendingSelfApply;
            pendingSelfApply = -1;
        }
        a.trailingCollides = groundOrColliders() ? trailingCollides : 0u;
        a.failure = failure.ptr;
        a.tileChanged = tileChanged;
        const bool leading = a.leadingCollide != 0 || a.selfAcc != nullptr || a.clearAcc != nullptr;
        a.prevChanged = leading || tileChanged == nullptr ? nullptr : prevChanged;
        a.tileNbrStart = t.nbrStart.ptr;
        a.tileNbrs = t.nbrs.ptr;
        a.roundStats = roundStats.ptr;
        // The odd program's tiles own the even program's knots, with a
        // smaller halo: the even program's windows serve both.
        const TileProgram& w = &t == &tileOdd ? tileEven : t;
        a.tileWindows = !scratch && tileWindowsBuilt && w.windows.count != 0 ? w.windows.ptr : nullptr;
        a.tileWindowSpheres = w.windowSpheres.ptr;
        for (uint32_t k = 0; k < t.passes; ++k) a.kinds[k] = t.kinds[k];
        const bool collides = a.leadingCollide != 0 || a.trailingCollides != 0;
        a.stageOffset = t.maxLoaded + (collides ? 2 * a.sphereCount : 0u);
        a.stageItems = t.staged ? t.maxItems : 0u;
        const size_t shared = a.stageOffset * sizeof(float4) + (t.staged ? stagedBytes(t.maxItems) : 0u);
        const bool strain = a.sweepPasses != 0;
        if (strain != t.packedStrain) fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "tile program item layout mismatch");
        if (t.staged) {
            if (strain) k_tile_program<true, true><<<t.tiles, t.block, shared, stream>>>(a);
            else k_tile_program<false, true><<<t.tiles, t.block, shared, stream>>>(a);
        } else {
            if (strain) k_tile_program<true, false><<<t.tiles, t.block, shared, stream>>>(a);
            else k_tile_program<false, false><<<t.tiles, t.block, shared, stream>>>(a);
        }
        if (scratch) return;
        posCur = a.posOut;
        if (t.hasDistance) dLamCur = a.distanceLambdaOut;
        if (t.hasBend) bLamCur = a.bendLambdaOut;
        if (t.hasKnot) kLamCur = a.knotLambdaOut;
    }

    // LDS of a staged program beyond the knots and spheres: 16-byte items and
    // their input multipliers, rounded to whole float4.
    static size_t stagedBytes(uint32_t items) {
        return static_cast<size_t>(items) * sizeof(uint4) + (static_cast<size_t>(items) * sizeof(float) + 15) / 16 * 16;
    }

    // Moves the live state back to the primary buffers at the end of a substep.
    void restorePrimaryBuffers() {
        settleSelf();
        if (posCur != pos.ptr) k_copy_float4<<<blocksFor(P), kBlock, 0, stream>>>(posCur, pos.ptr, P);
        if (D != 0 && dLamCur != dLambda.ptr) k_copy_float<<<blocksFor(D), kBlock, 0, stream>>>(dLamCur, dLambda.ptr, D);
        if (B != 0 && bLamCur != bLambda.ptr) k_copy_float<<<blocksFor(B), kBlock, 0, stream>>>(bLamCur, bLambda.ptr, B);
        if (K != 0 && kLamCur != kLambda.ptr) k_copy_float<<<blocksFor(K), kBlock, 0, stream>>>(kLamCur, kLambda.ptr, K);
        posCur = pos.ptr;
        dLamCur = dLambda.ptr;
        bLamCur = bLambda.ptr;
        kLamCur = kLambda.ptr;
    }

    // The substep's sphere windows of the even and strain programs' tiles,
    // from the positions before the first constraint pass (windowed scenes
    // with tile programs); the programs' ground/sphere passes then load them
    // instead of building their own.
    bool tileWindowsBuilt = false;
    void launchTileWindows() {
        tileWindowsBuilt = false;
        if (!useTiles || extension.staticSphereCount < sphereWindowMin || extension.staticSphereCount == 0) return;
        // Pending corrections may still move knots: a window needs no exact
        // box, only one that most knots stay inside.
        const uint32_t n = extension.staticSphereCount;
        for (TileProgram* t : {&tileEven, &tileStrain}) {
            if (!t->valid || t->windows.count == 0) continue;
            k_tile_windows<<<t->tiles, 1024, 2 * n * sizeof(float4), stream>>>(
                cfg(), posCur, t->loadStart.ptr, t->loadIds.ptr, spheres.ptr, n, t->windows.ptr, t->windowSpheres.ptr);
        }
        tileWindowsBuilt = tileEven.valid && tileEven.windows.count != 0 &&
                           (!tileStrain.valid || tileStrain.windows.count != 0);
    }

    // Strain sweeps followed by `collides` ground/sphere passes. flags >= 0:
    // the round's slot of per-tile change flags, which its strain program
    // records; repeat: the program repeats the previous slot's round with
    // nothing in between, so it skips the tiles whose loaded knots that round
    // left unchanged.
    void launchStrainAndCollide(uint32_t sweeps, uint32_t collides, int flags = -1, bool repeat = false) {
        if (useTiles && sweeps != 0 && tileStrain.valid) {  // consumes a deferred collide
            const size_t tiles = tileStrain.tiles;
            uint32_t* changed = flags >= 0 ? roundFlags.ptr + static_cast<size_t>(flags) * tiles : nullptr;
            const uint32_t* prev =
                flags > 0 && repeat ? roundFlags.ptr + static_cast<size_t>(flags - 1) * tiles : nullptr;
            launchTileProgram(tileStrain, collides, cfg(), ok(), false, changed, prev);
            return;
        }
        launchStrain(sweeps);
        for (uint32_t k = 0; k < collides; ++k) launchCollide();
    }

    void launchGrips() {
        if (G == 0) return;
        flushCollide();
        k_grip<<<blocksFor(G), kBlock, 0, stream>>>(cfg(), ok(), posCur, grips.ptr, G, P, failure.ptr);
    }

    // The substep in the reference's dispatch order
    // (tools/cloth_bag_metal.mm:4211-4709), families without records omitted.
    // substepConfigs / substepTimes: where k_prepare reads each substep's
    // config and end time (nullptr: the solver's own staging).
    // certify: run the certificate loop (host-synchronised, so never while
    // capturing) after the schedule's certificate rounds.
    void recordSubstep(uint32_t iterations, uint32_t sweeps, const NumiClothBagGPUConfig* substepConfigs = nullptr,
                       const double* substepTimes = nullptr, bool certify = false) {
        selfBuffer = 0;
        pendingSelfApply = -1;
        dirtySelf = -1;
        pendingCollide = false;
        posCur = pos.ptr;
        dLamCur = dLambda.ptr;
        bLamCur = bLambda.ptr;
        kLamCur = kLambda.ptr;
        k_prepare<<<1, 64, 0, stream>>>(substepConfigs != nullptr ? substepConfigs : configs.ptr,
                                        substepTimes != nullptr ? substepTimes : times.ptr, slot.ptr, failure.ptr);
        const float damping = extension.damping;
        // The damping factor depends on dt; dt is constant over a scene's
        // trajectories (the reference never changes it), so it is folded here.
        const float dt = initialConfig.gravityAndTimestep.w;
        const float factor = damping > 0.0f ? 1.0f / (1.0f + damping * dt) : 1.0f;
        const bool attach = G != 0;
        const bool separateAdvance = attach || yarnAir || F != 0;
        launchBegin(!separateAdvance, factor);
        if (F != 0) {
            k_fruit_begin<<<blocksFor(std::max({F, FP, Y})), kBlock, 0, stream>>>(
                cfg(), ok(), fruits.ptr, F, fruitPairs.ptr, FP, yarnContacts.ptr, Y, failure.ptr);
        }
        if (attach) {
            k_grip_attach<<<blocksFor(G), kBlock, 0, stream>>>(cfg(), ok(), posCur, grips.ptr, G, dCuff.ptr, P,
                                                               failure.ptr);
        }
        if (yarnAir) {
            k_aero_clear<<<blocksFor(P), kBlock, 0, stream>>>(ok(), aeroForces.ptr, P);
            for (const Range& r : distanceColours)
                k_aero_accumulate<<<blocksFor(r.count), kBlock, 0, stream>>>(cfg(), ok(), r.first, r.count, posCur,
                                                                             vel.ptr, dIJ.ptr, aeroForces.ptr,
                                                                             aeroStatus.ptr, failure.ptr);
            k_aero_reduce<<<1, 256, 0, stream>>>(cfg(), ok(), posCur, vel.ptr, aeroForces.ptr, P, aeroReduction.ptr);
            k_aero_apply<<<blocksFor(P), kBlock, 0, stream>>>(cfg(), ok(), posCur, vel.ptr, aeroForces.ptr,
                                                              aeroReduction.ptr, P, failure.ptr);
        }
        if (F != 0) k_fruit_aero<<<blocksFor(F), kBlock, 0, stream>>>(cfg(), ok(), fruits.ptr, F, aeroStatus.ptr,
                                                                      failure.ptr);
        if (separateAdvance) {
            k_advance<<<blocksFor(P), kBlock, 0, stream>>>(cfg(), ok(), posCur, vel.ptr, P);
            if (F != 0) k_fruit_advance<<<blocksFor(F), kBlock, 0, stream>>>(cfg(), ok(), fruits.ptr, F);
        }
        if (extension.sweptGround != 0u)
            k_swept_ground<<<blocksFor(std::max(P, F)), kBlock, 0, stream>>>(cfg(), ok(), posCur, prev.ptr, P,
                                                                              fruits.ptr, F);
        // Without grips or fruits nothing runs between a self-contact solve
        // and the next tile program: that program applies the corrections
        // (and the ground/sphere pass that follows in an iteration).
        const bool defer = self && !pointSelf && G == 0 && F == 0 && Y == 0 && useTiles;
        if (self) {
            launchBroadPhase(cfg(), ok());
            launchSelf(1, false, defer);
        }
        launchBuildYarnContacts();
        launchYarn(1);
        launchTileWindows();
        // Nothing between the constraint passes and the ground/sphere pass
        // of an iteration: the pass joins the tile program (or the
        // self-contact apply when that runs in between).
        const bool colliders = groundOrColliders();
        const bool trailing = colliders && !self && !pointSelf && G == 0 && F == 0;
        for (uint32_t it = 0; it < iterations; ++it) {
            TileProgram& program = it % 2 == 0 ? tileEven : tileOdd;
            const bool tiled = useTiles && program.valid;
            if (tiled) {
                launchTileProgram(program, trailing ? 1u : 0u, cfg(), ok());
            } else {
                launchDistances();
                launchKnots();
                if (it % 2 == 0) launchBends();
            }
            launchFruitPairsAndYarn(true, self && !pointSelf && selfInYarn);
            // The fruit ground pass joins the ground/sphere launch where one
            // follows here, or runs on its own.
            bool fruitGroundDone = false;
            if (pointSelf) {
                // segment self-contact, point self-collision, then ground
                if (self) launchSelf(0, false, false);
                launchPointSelf();
                if (colliders) {
                    launchCollide(true);
                    fruitGroundDone = true;
                }
            } else if (self) {
                launchSelf(0, colliders, defer, !defer);
                fruitGroundDone = !defer;
            } else if (!(tiled && trailing)) {
                launchCollide(true);
                fruitGroundDone = true;
            }
            if (!fruitGroundDone || !collideTakesFruitGround()) launchFruitGround();
        }
        launchStrainAndCollide(sweeps, 0);
        if (sweeps != 0) {
            // Reconciliation, endpoint and certificate passes
            // (reference tools/cloth_bag_metal.mm:4557-4590).
            // Without fruits nothing runs between the rounds' strain programs,
            // so a round after one that changed nothing is skipped (its
            // input and program are the unchanged round's).
            const bool repeating = F == 0 && FP == 0 && Y == 0;
            for (uint32_t pass = 0; pass < kReconciliationPasses; ++pass) {
                launchFruitPairsAndYarn();
                launchStrainAndCollide(sweeps, 1, repeating ? static_cast<int>(pass) : -1, pass > 0);
                launchFruitGround();
            }
            // Without self-contact the certificate rounds are the same strain
            // program again and again (their self-contact solve is none).
            const bool repeatingCertificates = repeating && !self && !pointSelf;
            for (uint32_t round = 0; round <= kCertificatePasses; ++round)
                certificateRound(sweeps, defer,
                                 repeatingCertificates ? static_cast<int>(kReconciliationPasses + round) : -1,
                                 round > 0);
            if (certify) runCertificates(sweeps, defer);
        }
        restorePrimaryBuffers();
        launchBuildYarnContacts();
        k_finalize<<<std::min(blocksFor(P), kFinalizeMaxBlocks), kBlock, 0, stream>>>(
            cfg(), ok(), pos.ptr, prev.ptr, vel.ptr, P, frictionStatus.ptr, failure.ptr);
        launchSelfFriction();
        if (F != 0) {
            k_fruit_finalize<<<blocksFor(F), kBlock, 0, stream>>>(cfg(), ok(), fruits.ptr, F, failure.ptr);
            if (Y != 0) {
                const size_t lds = static_cast<size_t>(P) * sizeof(float4) + static_cast<size_t>(D) * 4;
                k_yarn_friction_solve<<<1, kYarnThreads, lds, stream>>>(yarnArgs(0u, pos.ptr));
            }
            if (!pairColours.empty()) {
                k_fruit_pairs_friction<<<1, 64, 0, stream>>>(cfg(), ok(), pairColourRanges.ptr,
                                                             static_cast<uint32_t>(pairColours.size()), fruits.ptr, F,
                                                             fruitPairs.ptr, frictionStatus.ptr, failure.ptr);
            }
            k_fruit_ground_friction<<<blocksFor(F), kBlock, 0, stream>>>(cfg(), ok(), fruits.ptr, F,
                                                                         frictionStatus.ptr, failure.ptr);
            k_fruit_orientation<<<blocksFor(F), kBlock, 0, stream>>>(cfg(), ok(), fruits.ptr, F, failure.ptr);
            if (!rimHost.empty()) {
                k_release<<<1, fruitLanes(), 0, stream>>>(cfg(), ok(), pos.ptr, rim.ptr, fruits.ptr, F,
                                                          releaseStatus.ptr, mouthClearance.ptr, failure.ptr);
            }
        }
    }

    // One certificate round of the schedule: segment self-contact, the
    // strain sweeps, the final contact passes.
    // flags / repeat: as launchStrainAndCollide (scenes without fruits).
    void certificateRound(uint32_t sweeps, bool defer, int flags = -1, bool repeat = false) {
        launchSelf(0, false, defer);
        if (F == 0) {
            // Without fruits the final contact passes are ground passes only
            // and fuse into the strain program.
            launchStrainAndCollide(sweeps, kFinalContactPasses, flags, repeat);
        } else {
            launchStrainAndCollide(sweeps, 0);
            for (uint32_t pass = 0; pass < kFinalContactPasses; ++pass) launchFinalContacts();
        }
    }

    // The independent CPU reference's data-dependent certificate loop
    // (tools/cloth_bag.cpp:5143-5185): certificate rounds until the residuals
    // of one all fall below its thresholds, at most
    // extension.certificatePasses rounds. Each round's residuals are read
    // back (the stream is synchronised once per round).
    void runCertificates(uint32_t sweeps, bool defer) {
        certificateLast = 0;
        for (uint32_t pass = 0; pass < extension.certificatePasses; ++pass) {
            certificateRound(sweeps, defer);
            flushCollide();
            k_certificate_partial<<<certificateBlocks, kBlock, 0, stream>>>(
                cfg(), posCur, P, dIJ.ptr, dMaxLength.ptr, D, fruits.ptr, F, self ? grid.ptr : nullptr, pairs.ptr,
                certificatePartials.ptr);
            k_diag_final<<<1, kBlock, 0, stream>>>(certificatePartials.ptr, certificateBlocks, certificateResult.ptr);
            HIP_OK(hipMemcpyAsync(certificateHost, certificateResult.ptr, sizeof(float4), hipMemcpyDeviceToHost,
                                  stream));
            HIP_OK(hipStreamSynchronize(stream));
            certificateLast = pass + 1;
            const float4 r = *certificateHost;
            if (static_cast<double>(r.w) < 1.0e-6 && static_cast<double>(r.z) < 1.0e-6 &&
                static_cast<double>(r.y) < 1.0e-9 && static_cast<double>(r.x) < 1.0e-6) {
                break;
            }
        }
    }

    void ensureStrainProgram(uint32_t sweeps);

    void buildGraph(uint32_t iterations, uint32_t sweeps) {
        ensureStrainProgram(sweeps);
        if (!useGraphs) return;
        if (graph != nullptr && graphIterations == iterations && graphSweeps == sweeps) return;
        if (graph != nullptr) {
            HIP_OK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        recordSubstep(iterations, sweeps);
        hipGraph_t captured = nullptr;
        const hipError_t launchStatus = hipGetLastError();
        HIP_OK(hipStreamEndCapture(stream, &captured));
        if (launchStatus != hipSuccess) {
            (void)hipGraphDestroy(captured);
            fail(NUMI_CLOTH_HIP_ERROR_DEVICE, std::string("substep capture failed: ") + hipGetErrorString(launchStatus));
        }
        const hipError_t instantiate = hipGraphInstantiate(&graph, captured, nullptr, nullptr, 0);
        (void)hipGraphDestroy(captured);
        if (instantiate != hipSuccess) {
            graph = nullptr;
            fail(NUMI_CLOTH_HIP_ERROR_DEVICE, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(instantiate));
        }
        graphIterations = iterations;
        graphSweeps = sweeps;
    }

    void ensureStaging(uint32_t substeps) {
        if (substeps <= stagingCapacity) return;
        const uint32_t capacity = std::max(substeps, 64u);
        if (stagingConfigs != nullptr) HIP_OK(hipHostFree(stagingConfigs));
        if (stagingTimes != nullptr) HIP_OK(hipHostFree(stagingTimes));
        stagingConfigs = nullptr;
        stagingTimes = nullptr;
        HIP_OK(hipStreamSynchronize(stream));
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&stagingConfigs), capacity * sizeof(NumiClothBagGPUConfig)));
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&stagingTimes), capacity * sizeof(double)));
        configs.allocate(capacity);
        times.allocate(capacity);
        stagingCapacity = capacity;
        // The captured substep reads the per-substep arrays by address.
        if (graph != nullptr) {
            HIP_OK(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
    }
};

namespace {

void requireEqual(uint32_t a, uint32_t b, const char* what) {
    if (a != b) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("scene count mismatch: ") + what);
}

// Colour batches must tile their array in order, each record carrying the
// batch colour (the reference checks this per dispatch and raises
// FAILURE_BATCH, src/metal/ClothBag.metal:1172-1175; topology is immutable
// after create, so it is checked once here).
template <typename Colour>
std::vector<Range> checkBatches(const NumiClothBagGPUBatch* batches, uint32_t batchCount, uint32_t records,
                                Colour colourOf, const char* what) {
    std::vector<Range> ranges;
    if (batchCount != 0 && batches == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, std::string("null batches: ") + what);
    for (uint32_t b = 0; b < batchCount; ++b) {
        const mr_uint4 c = batches[b].control;
        if (static_cast<uint64_t>(c.x) + c.y > records) {
            fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("batch out of range (FAILURE_RANGE): ") + what);
        }
        for (uint32_t k = c.x; k < c.x + c.y; ++k) {
            if (colourOf(k) != c.z) {
                fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("record colour differs from its batch (FAILURE_BATCH): ") + what);
            }
        }
        if (c.y != 0) ranges.push_back({c.x, c.y});
    }
    return ranges;
}

// Within a colour no two records may share a particle; this is what makes a
// colour launch race-free (reference verifyColoring,
// tools/cloth_bag_metal.mm:1705-1819).
template <typename Participants>
void checkColouring(const std::vector<Range>& ranges, uint32_t particles, Participants participantsOf,
                    const char* what) {
    std::vector<uint32_t> stamp(particles, ~0u);
    for (uint32_t r = 0; r < ranges.size(); ++r) {
        for (uint32_t k = ranges[r].first; k < ranges[r].first + ranges[r].count; ++k) {
            uint32_t ids[4];
            const uint32_t n = participantsOf(k, ids);
            for (uint32_t m = 0; m < n; ++m) {
                if (ids[m] >= particles) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("particle index out of range (FAILURE_RANGE): ") + what);
                if (stamp[ids[m]] == r) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, std::string("colour batch shares a particle: ") + what);
                stamp[ids[m]] = r;
            }
        }
    }
}

void setupSelfContact(const NumiClothHipSceneDesc& scene, NumiClothHipSolver& s);
void setupPointSelf(NumiClothHipSolver& s);
void setupTiles(NumiClothHipSolver& s);
double timeTileProgram(NumiClothHipSolver& s, NumiClothHipSolver::TileProgram& t, uint32_t repeats);

void createSolver(const NumiClothHipSceneDesc& scene, NumiClothHipSolver& s) {
    if (scene.config == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "scene config is null");
    const NumiClothBagGPUConfig& c = *scene.config;
    if (c.control.x != NUMI_CLOTH_BAG_GPU_ABI_VERSION) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "config ABI version is not 13");
    requireEqual(c.control.y, scene.particleCount, "particles");
    requireEqual(c.control.z, scene.distanceCount, "distances");
    requireEqual(c.control.w, scene.gripCount, "grips");
    requireEqual(c.constraintCounts.x, scene.knotCount, "knots");
    requireEqual(c.constraintCounts.y, scene.bendCount, "bends");
    requireEqual(c.constraintCounts.w, scene.fruitCount, "fruits");
    requireEqual(c.contactCounts.x, scene.fruitPairCount, "fruit pairs");
    requireEqual(c.contactCounts.y, scene.yarnContactCount, "yarn contacts");
    if (scene.yarnContactCount != 0 && scene.yarnContactCount != scene.fruitCount * scene.distanceCount) {
        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "fruit/yarn records must be fruits x distances (FAILURE_RANGE)");
    }
    if (scene.fruitCount > 32) fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "at most 32 fruits (release masks are 32-bit)");
    s.initialConfig = c;
    s.P = scene.particleCount;
    s.D = scene.distanceCount;
    s.G = scene.gripCount;
    s.K = scene.knotCount;
    s.B = scene.bendCount;
    s.extension = scene.extension;
    s.driven = scene.extension.driveAmplitude != 0.0f || scene.extension.particleMotion != nullptr;
    // Fruit-only scenes (no knots) are valid, as in the reference's fruit probes.
    if (s.P == 0 && scene.fruitCount == 0) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "scene has no particles");
    if (s.P != 0 && scene.particles == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null particles");
    if (s.D != 0 && scene.distances == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null distances");
    if (s.B != 0 && scene.bends == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null bends");
    if (s.K != 0 && scene.knots == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null knots");
    if (s.G != 0 && scene.grips == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null grips");
    if (scene.extension.staticSphereCount > 4096) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "at most 4096 static spheres");

    s.hostDistances.assign(scene.distances, scene.distances + s.D);
    s.hostBends.assign(scene.bends, scene.bends + s.B);
    s.hostKnots.assign(scene.knots, scene.knots + s.K);
    s.distanceColours = checkBatches(scene.distanceBatches, scene.distanceBatchCount, s.D,
                                     [&](uint32_t k) { return s.hostDistances[k].particlesAndColor.z; }, "distance");
    s.bendColours = checkBatches(scene.bendBatches, scene.bendBatchCount, s.B,
                                 [&](uint32_t k) { return s.hostBends[k].particlesAndColor.w; }, "bend");
    s.knotColours = checkBatches(scene.knotBatches, scene.knotBatchCount, s.K,
                                 [&](uint32_t k) { return s.hostKnots[k].control.x; }, "knot");
    checkColouring(s.distanceColours, s.P, [&](uint32_t k, uint32_t* ids) {
        ids[0] = s.hostDistances[k].particlesAndColor.x;
        ids[1] = s.hostDistances[k].particlesAndColor.y;
        return 2u;
    }, "distance");
    checkColouring(s.bendColours, s.P, [&](uint32_t k, uint32_t* ids) {
        ids[0] = s.hostBends[k].particlesAndColor.x;
        ids[1] = s.hostBends[k].particlesAndColor.z;
        return 2u;
    }, "bend");
    checkColouring(s.knotColours, s.P, [&](uint32_t k, uint32_t* ids) {
        const mr_uint4 q = s.hostKnots[k].particles;
        ids[0] = q.x;
        ids[1] = q.y;
        ids[2] = q.z;
        ids[3] = q.w;
        return 4u;
    }, "knot");
    for (const NumiClothBagGPUBend& b : s.hostBends) {
        if (b.particlesAndColor.y >= s.P) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "bend middle index out of range");
    }
    if (scene.cuffCount != 0) {
        if (scene.cuffParticles == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null cuff");
        s.cuff.assign(scene.cuffParticles, scene.cuffParticles + scene.cuffCount);
        for (uint32_t p : s.cuff) {
            if (p >= s.P) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "cuff topology is out of range");
        }
    }

    HIP_OK(hipGetDevice(&s.device));
    HIP_OK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    hipStream_t st = s.stream;

    std::vector<float4> hp(s.P), hq(s.P), hv(s.P);
    for (uint32_t i = 0; i < s.P; ++i) {
        const NumiClothBagGPUParticle& p = scene.particles[i];
        hp[i] = make_float4(p.positionAndInverseMass.x, p.positionAndInverseMass.y, p.positionAndInverseMass.z,
                            p.positionAndInverseMass.w);
        hq[i] = make_float4(p.previousAndMass.x, p.previousAndMass.y, p.previousAndMass.z, p.previousAndMass.w);
        hv[i] = make_float4(p.velocity.x, p.velocity.y, p.velocity.z, p.velocity.w);
    }
    s.pos.allocate(s.P);
    s.prev.allocate(s.P);
    s.vel.allocate(s.P);
    s.pos.upload(hp.data(), s.P, st);
    s.prev.upload(hq.data(), s.P, st);
    s.vel.upload(hv.data(), s.P, st);
    if (scene.extension.particleMotion != nullptr) {
        std::vector<float4> motion(s.P);
        for (uint32_t i = 0; i < s.P; ++i) {
            const mr_float4 m = scene.extension.particleMotion[i];
            if (!std::isfinite(m.x) || !std::isfinite(m.y) || !std::isfinite(m.z) || !(m.w >= 0.0f) ||
                !std::isfinite(m.w)) {
                fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "particle motion must be finite with damping >= 0");
            }
            motion[i] = make_float4(m.x, m.y, m.z, m.w);
        }
        s.motion.allocate(s.P);
        s.motion.upload(motion.data(), s.P, st);
    }
    if (s.driven) {
        s.anchor.allocate(s.P);
        s.anchor.upload(hp.data(), s.P, st);
    }

    std::vector<uint2> ij(s.D);
    std::vector<float2> rc(s.D);
    std::vector<float> lam(s.D), maxLen(s.D);
    for (uint32_t k = 0; k < s.D; ++k) {
        const NumiClothBagGPUDistance& d = s.hostDistances[k];
        ij[k] = make_uint2(d.particlesAndColor.x, d.particlesAndColor.y);
        rc[k] = make_float2(d.material.x, d.material.y);
        lam[k] = d.material.z;
        maxLen[k] = d.material.x * (1.0f + d.material.w);
    }
    s.dIJ.allocate(s.D);
    s.dRestCompliance.allocate(s.D);
    s.dLambda.allocate(s.D);
    s.dMaxLength.allocate(s.D);
    s.dIJ.upload(ij.data(), s.D, st);
    s.dRestCompliance.upload(rc.data(), s.D, st);
    s.dLambda.upload(lam.data(), s.D, st);
    s.dMaxLength.upload(maxLen.data(), s.D, st);

    std::vector<uint2> ends(s.B);
    std::vector<float2> cc(s.B);
    std::vector<float> bl(s.B);
    for (uint32_t k = 0; k < s.B; ++k) {
        const NumiClothBagGPUBend& b = s.hostBends[k];
        ends[k] = make_uint2(b.particlesAndColor.x, b.particlesAndColor.z);
        cc[k] = make_float2(b.material.x, b.material.z);
        bl[k] = b.material.w;
    }
    s.bEnds.allocate(s.B);
    s.bChordCompliance.allocate(s.B);
    s.bLambda.allocate(s.B);
    s.bEnds.upload(ends.data(), s.B, st);
    s.bChordCompliance.upload(cc.data(), s.B, st);
    s.bLambda.upload(bl.data(), s.B, st);

    std::vector<uint4> quad(s.K);
    std::vector<float2> kc(s.K);
    std::vector<float> kl(s.K);
    for (uint32_t k = 0; k < s.K; ++k) {
        const NumiClothBagGPUKnot& n = s.hostKnots[k];
        quad[k] = make_uint4(n.particles.x, n.particles.y, n.particles.z, n.particles.w);
        kc[k] = make_float2(n.material.x, n.material.y);
        kl[k] = n.material.z;
    }
    s.kQuad.allocate(s.K);
    s.kCosineCompliance.allocate(s.K);
    s.kLambda.allocate(s.K);
    s.kQuad.upload(quad.data(), s.K, st);
    s.kCosineCompliance.upload(kc.data(), s.K, st);
    s.kLambda.upload(kl.data(), s.K, st);

    s.grips.allocate(s.G);
    s.grips.upload(scene.grips, s.G, st);
    s.F = scene.fruitCount;
    s.FP = scene.fruitPairCount;
    s.Y = scene.yarnContactCount;
    if (s.F != 0) {
        if (scene.fruits == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null fruits");
        s.fruits.allocate(s.F);
        s.fruits.upload(scene.fruits, s.F, st);
        std::vector<NumiClothBagGPUFruitPair> pairs(scene.fruitPairs, scene.fruitPairs + s.FP);
        s.pairColours = checkBatches(scene.fruitPairBatches, scene.fruitPairBatchCount, s.FP,
                                     [&](uint32_t k) { return pairs[k].fruitsAndColor.z; }, "fruit pair");
        std::vector<uint32_t> pairStamp(s.F, ~0u);
        for (uint32_t r = 0; r < s.pairColours.size(); ++r) {
            for (uint32_t k = s.pairColours[r].first; k < s.pairColours[r].first + s.pairColours[r].count; ++k) {
                const mr_uint4 p = pairs[k].fruitsAndColor;
                if (p.x >= s.F || p.y >= s.F || p.x == p.y) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "fruit pair index (FAILURE_RANGE)");
                if (pairStamp[p.x] == r || pairStamp[p.y] == r) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "fruit pair colour shares a fruit");
                pairStamp[p.x] = r;
                pairStamp[p.y] = r;
            }
        }
        s.fruitPairs.allocate(s.FP);
        s.fruitPairs.upload(pairs.data(), s.FP, st);
        std::vector<uint2> ranges;
        std::vector<uint32_t> colourOf(std::max(s.FP, 1u), ~0u);
        for (uint32_t r = 0; r < s.pairColours.size(); ++r) {
            ranges.push_back(make_uint2(s.pairColours[r].first, s.pairColours[r].count));
            for (uint32_t k = 0; k < s.pairColours[r].count; ++k) colourOf[s.pairColours[r].first + k] = r;
        }
        s.pairColourRanges.allocate(ranges.size());
        s.pairColourRanges.upload(ranges.data(), ranges.size(), st);
        s.pairColourOf.allocate(colourOf.size());
        s.pairColourOf.upload(colourOf.data(), colourOf.size(), st);
        s.yarnContacts.allocate(s.Y);
        s.yarnContacts.upload(scene.yarnContacts, s.Y, st);
        if (s.Y != 0) {
            if (s.F > 32) fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "fruit/yarn solve runs one fruit per lane of one wave");
            int maxShared = 0;
            HIP_OK(hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, s.device));
            // Static LDS of the yarn solve (ranking scratch, candidate list,
            // staged visits, fruits) plus the knots, their travel and the
            // per-segment fruit masks.
            hipFuncAttributes attributes{};
            HIP_OK(hipFuncGetAttributes(&attributes, reinterpret_cast<const void*>(&k_yarn_solve)));
            const uint64_t fixed = attributes.sharedSizeBytes;
            const uint64_t dynamic = static_cast<uint64_t>(s.P) * 20 + static_cast<uint64_t>(s.D) * 8;
            if (fixed + dynamic > static_cast<uint64_t>(maxShared)) {
                fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "fruit/yarn contact stages every knot in LDS");
            }
            std::vector<uint4> batches;
            std::vector<uint2> segmentBatch(s.D, make_uint2(0u, 0u));
            for (const Range& r : s.distanceColours) {
                batches.push_back(make_uint4(r.first, r.count, 0u, 0u));
                for (uint32_t k = r.first; k < r.first + r.count; ++k) segmentBatch[k] = make_uint2(r.first, r.count);
            }
            s.distanceBatchesDevice.allocate(batches.size());
            s.distanceBatchesDevice.upload(batches.data(), batches.size(), st);
            s.segmentBatch.allocate(s.D);
            s.segmentBatch.upload(segmentBatch.data(), s.D, st);
            s.yarnStats.allocate(4);
            HIP_OK(hipMemsetAsync(s.yarnStats.ptr, 0, 4 * sizeof(uint32_t), st));
            if (const char* env = std::getenv("NUMI_CLOTH_HIP_YARN_LIST")) s.yarnList = std::strcmp(env, "0") != 0;
            if (const char* env = std::getenv("NUMI_CLOTH_HIP_YARN_VISIT_BY_VISIT"))
                s.yarnVisitByVisit = std::strcmp(env, "0") != 0;
            // One workgroup walks every self-contact candidate: bag-sized
            // scenes only.
            s.selfInYarn = s.P <= 8192u;
            if (const char* env = std::getenv("NUMI_CLOTH_HIP_SELF_IN_YARN"))
                s.selfInYarn = s.selfInYarn && std::strcmp(env, "0") != 0;
            s.yarnListHeader.allocate(4);
            s.yarnListPairs.allocate(kYarnListCapacity);
            s.yarnListKnots.allocate(s.P);
            s.yarnListFruits.allocate(32);
            HIP_OK(hipMemsetAsync(s.yarnListHeader.ptr, 0, 4 * sizeof(uint32_t), st));
            HIP_OK(hipMemsetAsync(s.yarnListKnots.ptr, 0, s.P * sizeof(float4), st));
            HIP_OK(hipMemsetAsync(s.yarnListFruits.ptr, 0, 32 * sizeof(float4), st));
        }
        if (scene.mouthRimCount != 0) {
            if (scene.mouthRimParticles == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null mouth rim");
            s.rimHost.assign(scene.mouthRimParticles, scene.mouthRimParticles + scene.mouthRimCount);
            for (uint32_t p : s.rimHost)
                if (p >= s.P) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "mouth rim index (FAILURE_RANGE)");
            if (c.mouthControl.x > s.rimHost.size()) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "mouth rim shorter than mouthControl.x");
            s.rim.allocate(s.rimHost.size());
            s.rim.upload(s.rimHost.data(), s.rimHost.size(), st);
        }
    }
    s.yarnAir = c.airVelocityAndDensity.w > 0.0f &&
                (c.aerodynamicCoefficients.x > 0.0f || c.aerodynamicCoefficients.y > 0.0f) && s.D != 0;
    if (s.yarnAir) {
        s.aeroForces.allocate(s.P);
        s.aeroReduction.allocate(1);
    }
    s.mouthClearance.allocate(32);
    HIP_OK(hipMemsetAsync(s.mouthClearance.ptr, 0, 32 * sizeof(uint32_t), st));
    s.dCuff.allocate(s.cuff.size());
    s.dCuff.upload(s.cuff.data(), s.cuff.size(), st);
    const uint32_t nSpheres = scene.extension.staticSphereCount;
    s.spheres.allocate(nSpheres);
    if (nSpheres != 0) {
        std::vector<float4> sp(nSpheres);
        for (uint32_t k = 0; k < nSpheres; ++k) {
            const mr_float4& v = scene.extension.staticSpheres[k];
            sp[k] = make_float4(v.x, v.y, v.z, v.w);
        }
        s.spheres.upload(sp.data(), nSpheres, st);
        HIP_OK(hipStreamSynchronize(st));
    }
    s.extension.staticSpheres = nullptr;  // never dereference caller memory later

    s.slot.allocate(1);
    s.failure.allocate(1);
    s.roundStats.allocate(2);
    HIP_OK(hipMemsetAsync(s.roundStats.ptr, 0, 2 * sizeof(unsigned long long), st));
    s.selfStatus.allocate(1);
    s.frictionStatus.allocate(1);
    s.aeroStatus.allocate(1);
    s.releaseStatus.allocate(1);
    HIP_OK(hipMemsetAsync(s.slot.ptr, 0, sizeof(SubstepSlot), st));
    HIP_OK(hipMemsetAsync(s.failure.ptr, 0, sizeof(uint32_t), st));
    HIP_OK(hipMemsetAsync(s.selfStatus.ptr, 0, sizeof(NumiClothBagGPUSelfStatus), st));
    HIP_OK(hipMemsetAsync(s.frictionStatus.ptr, 0, sizeof(NumiClothBagGPUFrictionStatus), st));
    HIP_OK(hipMemsetAsync(s.aeroStatus.ptr, 0, sizeof(NumiClothBagGPUAerodynamicsStatus), st));
    // Release masks carry over between calls (reference
    // tools/cloth_bag_metal.mm:4161-4162, :4779-4780).
    const NumiClothBagGPUReleaseStatus masks{{scene.initialCandidateMask, scene.initialReleasedMask, 0u, 0u}};
    s.releaseStatus.upload(&masks, 1, st);
    s.diagPartials.allocate(1024);
    if (const char* graphs = std::getenv("NUMI_CLOTH_HIP_GRAPHS")) {
        s.useGraphs = std::strcmp(graphs, "0") != 0;
        s.launchDecided = true;
    }
    // "0": test every sphere list whole; "1": window every non-empty list.
    if (const char* window = std::getenv("NUMI_CLOTH_HIP_SPHERE_WINDOW"))
        s.sphereWindowMin = std::strcmp(window, "0") == 0 ? UINT32_MAX : 1u;
    s.posCur = s.pos.ptr;
    s.dLamCur = s.dLambda.ptr;
    s.bLamCur = s.bLambda.ptr;
    s.kLamCur = s.kLambda.ptr;
    s.debugSlot.allocate(1);
    s.initialPositions = hp;
    if (scene.extension.selfContact != 0) setupSelfContact(scene, s);
    if (scene.extension.pointSelfCollision != 0) setupPointSelf(s);
    bool tiles = true;
    if (const char* env = std::getenv("NUMI_CLOTH_HIP_TILES")) tiles = std::strcmp(env, "0") != 0;
    if (tiles) setupTiles(s);
    // The strain-limit program of the reference's default 3 sweeps, so that
    // numi_cloth_hip_step_noalloc can run them without building one.
    s.ensureStrainProgram(3);
    s.diagResult.allocate(1);
    if (scene.extension.certificatePasses != 0) {
        const uint64_t work = std::max<uint64_t>({s.P, s.D, static_cast<uint64_t>(s.F) * s.D, s.pairCapacity});
        s.certificateBlocks = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(1, (work + kBlock - 1) / kBlock)));
        s.certificatePartials.allocate(s.certificateBlocks);
        s.certificateResult.allocate(1);
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&s.certificateHost), sizeof(float4)));
    }
    s.ensureStaging(64);
    HIP_OK(hipStreamSynchronize(st));
}

// Two-hop neighbourhoods over the distance graph (sorted CSR, each list
// includes its own knot): the locality rule that keeps adjacent yarn out of
// self-contact (reference tools/cloth_bag_metal.mm:1107-1152).
void twoHopLists(const NumiClothHipSolver& s, std::vector<uint32_t>& start, std::vector<uint32_t>& list) {
    std::vector<std::vector<uint32_t>> direct(s.P);
    for (const NumiClothBagGPUDistance& d : s.hostDistances) {
        direct[d.particlesAndColor.x].push_back(d.particlesAndColor.y);
        direct[d.particlesAndColor.y].push_back(d.particlesAndColor.x);
    }
    start.assign(s.P + 1, 0);
    list.clear();
    for (uint32_t i = 0; i < s.P; ++i) {
        std::vector<uint32_t> near{i};
        for (uint32_t a : direct[i]) {
            near.push_back(a);
            for (uint32_t b : direct[a]) near.push_back(b);
        }
        std::sort(near.begin(), near.end());
        near.erase(std::unique(near.begin(), near.end()), near.end());
        start[i + 1] = start[i] + static_cast<uint32_t>(near.size());
        list.insert(list.end(), near.begin(), near.end());
    }
}

// Point self-collision buffers: a power-of-two bucket table of at least
// twice the knots, the two-hop lists (shared with segment self-contact).
void setupPointSelf(NumiClothHipSolver& s) {
    if (s.P == 0) return;
    hipStream_t st = s.stream;
    if (s.localStart.count == 0) {
        std::vector<uint32_t> start, list;
        twoHopLists(s, start, list);
        s.localStart.allocate(start.size());
        s.localStart.upload(start.data(), start.size(), st);
        s.localList.allocate(std::max<size_t>(list.size(), 1));
        s.localList.upload(list.data(), list.size(), st);
    }
    uint32_t buckets = 1024;
    while (buckets < 2u * s.P) buckets <<= 1;
    s.pointBuckets = buckets;
    s.pointEnds.allocate(buckets);
    s.pointIds.allocate(s.P);
    s.pointAcc.allocate(s.P);
    HIP_OK(hipMemsetAsync(s.pointAcc.ptr, 0, s.P * sizeof(FixedAccumulator), st));
    s.pointTouched.allocate((s.P + 31) / 32);
    HIP_OK(hipMemsetAsync(s.pointTouched.ptr, 0, (s.P + 31) / 32 * sizeof(uint32_t), st));
    s.pointContacts.allocate(1);
    HIP_OK(hipMemsetAsync(s.pointContacts.ptr, 0, sizeof(uint32_t), st));
    s.pointBucketCount.allocate(1);
    s.pointBucketCount.upload(&buckets, 1, st);
    s.pointScanSums.allocate((buckets + kScanTile - 1) / kScanTile + 1);
    s.pointTotal.allocate(1);
    s.pointSelf = true;
}

void setupSelfContact(const NumiClothHipSceneDesc& scene, NumiClothHipSolver& s) {
    if (s.D == 0) return;
    if (!(scene.config->clothMaterial.y > 0.0f)) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "self-contact needs a positive cell size (clothMaterial.y)");
    }
    std::vector<uint32_t> start, list;
    twoHopLists(s, start, list);
    hipStream_t st = s.stream;
    if (s.D > kEntrySegmentMask)
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "hash self-contact supports at most 2^29 - 1 segments");
    s.self = true;
    s.localStart.allocate(start.size());
    s.localStart.upload(start.data(), start.size(), st);
    s.localList.allocate(list.size());
    s.localList.upload(list.data(), list.size(), st);
    uint32_t longest = 0;
    for (uint32_t i = 0; i < s.P; ++i) longest = std::max(longest, start[i + 1] - start[i]);
    s.localWidth = longest <= 16 ? 16u : (longest <= 32 ? 32u : 0u);
    if (s.localWidth != 0) {
        std::vector<uint32_t> fixed(static_cast<size_t>(s.P) * s.localWidth, ~0u);
        for (uint32_t i = 0; i < s.P; ++i)
            for (uint32_t k = start[i]; k < start[i + 1]; ++k) fixed[static_cast<size_t>(i) * s.localWidth + k - start[i]] = list[k];
        s.localFixed.allocate(fixed.size());
        s.localFixed.upload(fixed.data(), fixed.size(), st);
    }
    // 32 cell entries per segment: a segment whose swept box (start and
    // predicted ends, grown by radius + margin) spans up to 4 x 4 x 2 cells.
    // More is a RANGE failure, as is a candidate list beyond its capacity.
    const uint64_t entries = static_cast<uint64_t>(s.D) * 32u;
    const uint64_t pairCapacity = scene.extension.selfCandidateCapacity != 0
                                      ? scene.extension.selfCandidateCapacity
                                      : static_cast<uint64_t>(s.D) * 16u;
    if (entries > 0x7fffffffull || pairCapacity > 0x7fffffffull) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "self-contact capacities exceed 2^31");
    }
    s.entryCapacity = static_cast<uint32_t>(entries);
    s.pairCapacity = static_cast<uint32_t>(pairCapacity);
    s.entryTiles = (s.entryCapacity + kScanTile - 1) / kScanTile;
    s.segmentBlocks = (s.D + kScanThreads - 1) / kScanThreads;
    // Grids of up to max(entry capacity, 2^22) cells sort their entries by
    // counting; larger (sparse) grids by radix. NUMI_CLOTH_HIP_SELF_SORT=radix
    // sorts every grid by radix.
    const char* sortMode = std::getenv("NUMI_CLOTH_HIP_SELF_SORT");
    const bool radixOnly = sortMode != nullptr && std::strcmp(sortMode, "radix") == 0;
    s.cellCapacity = radixOnly ? 0u : std::max(s.entryCapacity, 1u << 22);
    if (s.cellCapacity != 0) s.cellSlots.allocate(s.cellCapacity);
    s.boxLo.allocate(s.D);
    s.boxHi.allocate(s.D);
    s.partialLo.allocate(s.segmentBlocks);
    s.partialHi.allocate(s.segmentBlocks);
    s.grid.allocate(1);
    s.segmentCounts.allocate(s.D);
    s.entryCounts.allocate(s.entryCapacity);
    const uint32_t maxScanLength = std::max({s.D, kRadixBins * s.entryTiles, s.entryCapacity, s.cellCapacity, 1u});
    s.scanSums.allocate((maxScanLength + kScanTile - 1) / kScanTile + 1);
    s.totals.allocate(4);
    const uint32_t histLength = kRadixBins * s.entryTiles;
    s.keysA.allocate(s.entryCapacity);
    s.valsA.allocate(s.entryCapacity);
    s.keysB.allocate(s.entryCapacity);
    s.valsB.allocate(s.entryCapacity);
    s.histogram.allocate(histLength);
    s.pairs.allocate(s.pairCapacity);
    s.pairKnots.allocate(s.pairCapacity);
    if (s.localWidth != 0) s.foundPairs.allocate(s.pairCapacity);
    s.selfImpulses.allocate(s.pairCapacity);  // written by the emit pass before any use
    for (int b = 0; b < 2; ++b) {
        s.accumulator[b].allocate(s.P);
        HIP_OK(hipMemsetAsync(s.accumulator[b].ptr, 0, s.P * sizeof(FixedAccumulator), st));
        s.touched[b].allocate((s.P + 31) / 32);
        HIP_OK(hipMemsetAsync(s.touched[b].ptr, 0, (s.P + 31) / 32 * sizeof(uint32_t), st));
    }
    HIP_OK(hipMemsetAsync(s.grid.ptr, 0, sizeof(SelfGrid), st));
    HIP_OK(hipStreamSynchronize(st));
}

// ---- tile programs -------------------------------------------------------

struct TileConstants {
    std::vector<float> distanceRest, distanceCompliance, distanceMaxLength, bendRest, bendCompliance;
};

TileTopology tileTopology(const NumiClothHipSolver& s, std::vector<uint32_t>& dPairs, std::vector<uint32_t>& bPairs,
                          std::vector<uint32_t>& kQuads, TileConstants& k) {
    k.distanceRest.resize(s.D);
    k.distanceCompliance.resize(s.D);
    k.distanceMaxLength.resize(s.D);
    for (uint32_t c = 0; c < s.D; ++c) {
        const mr_float4 m = s.hostDistances[c].material;
        k.distanceRest[c] = m.x;
        k.distanceCompliance[c] = m.y;
        k.distanceMaxLength[c] = m.x * (1.0f + m.w);  // the value uploaded as dMaxLength
    }
    k.bendRest.resize(s.B);
    k.bendCompliance.resize(s.B);
    for (uint32_t c = 0; c < s.B; ++c) {
        k.bendRest[c] = s.hostBends[c].material.x;
        k.bendCompliance[c] = s.hostBends[c].material.z;
    }
    dPairs.resize(2 * s.D);
    for (uint32_t k = 0; k < s.D; ++k) {
        dPairs[2 * k] = s.hostDistances[k].particlesAndColor.x;
        dPairs[2 * k + 1] = s.hostDistances[k].particlesAndColor.y;
    }
    bPairs.resize(2 * s.B);
    for (uint32_t k = 0; k < s.B; ++k) {
        bPairs[2 * k] = s.hostBends[k].particlesAndColor.x;
        bPairs[2 * k + 1] = s.hostBends[k].particlesAndColor.z;
    }
    kQuads.resize(4 * s.K);
    for (uint32_t k = 0; k < s.K; ++k) {
        const mr_uint4 q = s.hostKnots[k].particles;
        kQuads[4 * k] = q.x;
        kQuads[4 * k + 1] = q.y;
        kQuads[4 * k + 2] = q.z;
        kQuads[4 * k + 3] = q.w;
    }
    TileTopology t;
    t.particles = s.P;
    t.distancePairs = dPairs.data();
    t.bendPairs = bPairs.data();
    t.knotQuads = kQuads.data();
    t.positions = reinterpret_cast<const float*>(s.initialPositions.data());
    t.distanceRest = k.distanceRest.data();
    t.distanceCompliance = k.distanceCompliance.data();
    t.distanceMaxLength = k.distanceMaxLength.data();
    t.bendRest = k.bendRest.data();
    t.bendCompliance = k.bendCompliance.data();
    return t;
}

// The passes of one tile program: an iteration (distance, knot and, on even
// iterations, bend colours, in the reference's order) or a strain-limit
// block of `strainSweeps` sweeps over the distance colours.
std::vector<TilePass> tilePasses(const NumiClothHipSolver& s, bool bends, uint32_t strainSweeps) {
    std::vector<TilePass> passes;
    if (strainSweeps != 0) {
        for (uint32_t sweep = 0; sweep < strainSweeps; ++sweep)
            for (const Range& r : s.distanceColours) passes.push_back({kTileStrain, r.first, r.count});
        return passes;
    }
    for (const Range& r : s.distanceColours) passes.push_back({kTileDistance, r.first, r.count});
    for (const Range& r : s.knotColours) passes.push_back({kTileKnot, r.first, r.count});
    if (bends)
        for (const Range& r : s.bendColours) passes.push_back({kTileBend, r.first, r.count});
    return passes;
}

bool buildTileProgram(NumiClothHipSolver& s, const std::vector<TilePass>& passes, uint32_t owned,
                      NumiClothHipSolver::TileProgram& out) {
    out = NumiClothHipSolver::TileProgram{};
    if (passes.empty() || passes.size() > kMaxTilePasses) return false;
    std::vector<uint32_t> dPairs, bPairs, kQuads;
    TileConstants constants;
    const TileTopology topo = tileTopology(s, dPairs, bPairs, kQuads, constants);
    TilePlan plan;
    if (!planTiles(topo, passes, owned, s.tileCapacity, 32, &plan)) return false;
    out.valid = true;
    out.tiles = plan.tiles;
    out.passes = static_cast<uint32_t>(passes.size());
    out.maxLoaded = plan.maxLoaded;
    out.owned = owned;
    out.halo = plan.halo;
    out.computed = plan.computed;
    out.unique = plan.unique;
    out.block = s.tileBlock;
    double bytes = s.P * (16.0 + 12.0);
    bool strain = false;
    for (const TilePass& pass : passes) {
        out.kinds.push_back(pass.kind);
        out.hasDistance |= pass.kind == kTileDistance;
        out.hasBend |= pass.kind == kTileBend;
        out.hasKnot |= pass.kind == kTileKnot;
        strain |= pass.kind == kTileStrain;
    }
    // Unique bytes one launch must move: knots read (xyzw) and written (xyz)
    // once, each constraint's static record and multiplier once.
    if (out.hasDistance) bytes += s.D * (20.0 + 8.0);
    if (out.hasBend) bytes += s.B * (16.0 + 8.0);
    if (out.hasKnot) bytes += s.K * (24.0 + 8.0);
    if (strain) bytes += s.D * 12.0;
    out.algorithmicBytes = bytes;
    hipStream_t st = s.stream;
    out.loadStart.allocate(plan.loadStart.size());
    out.loadStart.upload(plan.loadStart.data(), plan.loadStart.size(), st);
    out.loadIds.allocate(plan.loadIds.size());
    out.loadIds.upload(plan.loadIds.data(), plan.loadIds.size(), st);
    out.ownedCount.allocate(plan.ownedCount.size());
    out.ownedCount.upload(plan.ownedCount.data(), plan.ownedCount.size(), st);
    out.itemStart.allocate(plan.itemStart.size());
    out.itemStart.upload(plan.itemStart.data(), plan.itemStart.size(), st);
    for (uint32_t tile = 0; tile < plan.tiles; ++tile)
        out.maxItems = std::max(out.maxItems, plan.itemStart[(tile + 1) * out.passes] - plan.itemStart[tile * out.passes]);
    {
        std::vector<uint32_t> ownerOf(s.P, 0u);
        for (uint32_t tile = 0; tile < plan.tiles; ++tile)
            for (uint32_t k = 0; k < plan.ownedCount[tile]; ++k) ownerOf[plan.loadIds[plan.loadStart[tile] + k]] = tile;
        std::vector<uint32_t> start{0u}, list, seen;
        for (uint32_t tile = 0; tile < plan.tiles; ++tile) {
            seen.clear();
            for (uint32_t k = plan.loadStart[tile]; k < plan.loadStart[tile + 1]; ++k) seen.push_back(ownerOf[plan.loadIds[k]]);
            std::sort(seen.begin(), seen.end());
            seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
            list.insert(list.end(), seen.begin(), seen.end());
            start.push_back(static_cast<uint32_t>(list.size()));
        }
        out.nbrStart.allocate(start.size());
        out.nbrStart.upload(start.data(), start.size(), st);
        out.nbrs.allocate(list.size());
        out.nbrs.upload(list.data(), list.size(), st);
    }
    const uint32_t spheres = s.extension.staticSphereCount;
    if (spheres != 0 && spheres >= s.sphereWindowMin) {
        out.windows.allocate(out.tiles);
        out.windowSpheres.allocate(static_cast<size_t>(out.tiles) * spheres);
    }
    // Staged: knots, both sphere lists and the whole program in LDS.
    const uint64_t stagedLds = (static_cast<uint64_t>(out.maxLoaded) + 2ull * s.extension.staticSphereCount) * 16u +
                               NumiClothHipSolver::stagedBytes(out.maxItems);
    out.stageable = stagedLds <= s.ldsBytes;
    out.staged = out.stageable && s.tileStaged;
    if (strain) {
        // A strain program reads 8-byte items (local endpoints, rest * (1 +
        // limit)), two per uint4; its passes are all strain passes.
        for (const TilePass& pass : passes)
            if (pass.kind != kTileStrain && pass.kind != kTileCollide)
                fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "a strain program mixes pass kinds");
        const size_t n = plan.items.size() / 4;
        std::vector<uint32_t> packed((n + 1) / 2 * 4, 0u);
        for (size_t j = 0; j < n; ++j) {
            packed[2 * j] = plan.items[4 * j];
            packed[2 * j + 1] = plan.items[4 * j + 2];
        }
        out.packedStrain = true;
        out.items.allocate(packed.size() / 4);
        out.items.upload(reinterpret_cast<const uint4*>(packed.data()), packed.size() / 4, st);
    } else {
        out.items.allocate(plan.items.size() / 4);
        out.items.upload(reinterpret_cast<const uint4*>(plan.items.data()), plan.items.size() / 4, st);
    }
    HIP_OK(hipStreamSynchronize(st));
    return true;
}

// Launch time of one tile program writing to the alternate buffers only, so
// the state is not changed.
double timeTileProgram(NumiClothHipSolver& s, NumiClothHipSolver::TileProgram& t, uint32_t repeats) {
    hipEvent_t start = nullptr, stop = nullptr;
    HIP_OK(hipEventCreate(&start));
    HIP_OK(hipEventCreate(&stop));
    const NumiClothBagGPUConfig* config = &s.debugSlot.ptr->config;
    const uint32_t* valid = &s.debugSlot.ptr->ok;
    s.launchTileProgram(t, 0, config, valid, true);
    HIP_OK(hipEventRecord(start, s.stream));
    for (uint32_t k = 0; k < repeats; ++k) s.launchTileProgram(t, 0, config, valid, true);
    HIP_OK(hipEventRecord(stop, s.stream));
    HIP_OK(hipEventSynchronize(stop));
    HIP_OK(hipGetLastError());
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    return ms / repeats;
}

// Autotuner: picks the tile size (owned knots per workgroup) and the
// workgroup size by timing the even-iteration program, the most frequent
// launch of a substep, over a small candidate grid.
void setupTiles(NumiClothHipSolver& s) {
    if (s.D == 0 && s.B == 0 && s.K == 0) return;
    int maxShared = 0;
    HIP_OK(hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, s.device));
    const uint32_t spheres = s.extension.staticSphereCount;
    // Dynamic LDS: the loaded knots, the spheres and a sphere window's list;
    // 2 KiB stay free for the kernel's static LDS (sphere window, reduction
    // scratch, pass starts and kinds: 1,124 bytes).
    const uint64_t reserved = kTileStaticLds + 2ull * spheres * sizeof(float4);
    if (reserved + 64 * sizeof(float4) >= static_cast<uint64_t>(maxShared)) return;
    s.tileCapacity = static_cast<uint32_t>((static_cast<uint64_t>(maxShared) - reserved) / sizeof(float4));
    s.ldsBytes = static_cast<uint64_t>(maxShared) - kTileStaticLds;
    s.posAlt.allocate(s.P);
    s.dLambdaAlt.allocate(s.D);
    s.bLambdaAlt.allocate(s.B);
    s.kLambdaAlt.allocate(s.K);
    SubstepSlot slot{};
    slot.config = s.initialConfig;
    slot.ok = 1u;
    s.debugSlot.upload(&slot, 1, s.stream);
    const std::vector<TilePass> even = tilePasses(s, true, 0);
    double best = 1e30;
    uint32_t bestOwned = 0, bestBlock = 256;
    bool bestStaged = false;
    std::vector<uint32_t> sizes;
    if (s.P <= 4096) sizes.push_back(s.P);
    // Smaller tiles only add halo work once the tiles of 4096 knots already
    // give every CU a workgroup, so large scenes try fewer sizes; scenes of
    // at most 256 x 1024 knots also try tiles of 128 and 64 knots, which put
    // several workgroups on each CU to hide the projections' latency.
    const uint32_t smallest = s.P >= 1024u * 4096u ? 2048u
                              : (s.P >= 256u * 4096u ? 1024u : (s.P > 256u * 1024u ? 256u : 64u));
    for (uint32_t owned : {4096u, 2048u, 1024u, 512u, 256u, 128u, 64u})
        if (owned < s.P && owned >= smallest) sizes.push_back(owned);
    // Pins of the autotuner's axes, so that separate profiler runs measure one
    // configuration: NUMI_CLOTH_HIP_TILE_OWNED / _BLOCK name a candidate (the
    // tiles are off when the pinned one cannot be built);
    // NUMI_CLOTH_HIP_TILE_STAGED=0/1 overrides staging after the choice.
    auto pin = [](const char* name) -> long {
        const char* env = std::getenv(name);
        return env != nullptr ? std::strtol(env, nullptr, 10) : -1;
    };
    const long pinOwned = pin("NUMI_CLOTH_HIP_TILE_OWNED"), pinBlock = pin("NUMI_CLOTH_HIP_TILE_BLOCK");
    for (uint32_t owned : sizes) {
        if (pinOwned >= 0 && owned != static_cast<uint32_t>(pinOwned)) continue;
        NumiClothHipSolver::TileProgram candidate;
        if (!buildTileProgram(s, even, owned, candidate)) continue;
        for (uint32_t block : {256u, 512u, 1024u})
            for (bool staged : {false, true}) {
                if (pinBlock >= 0 && block != static_cast<uint32_t>(pinBlock)) continue;
                if (staged && !candidate.stageable) continue;
                candidate.block = block;
                candidate.staged = staged;
                const double ms = timeTileProgram(s, candidate, 5);
                if (ms < best) {
                    best = ms;
                    bestOwned = owned;
                    bestBlock = block;
                    bestStaged = staged;
                }
            }
    }
    if (bestOwned == 0) return;
    s.tileOwned = bestOwned;
    s.tileBlock = bestBlock;
    s.tileStaged = bestStaged;
    if (const char* env = std::getenv("NUMI_CLOTH_HIP_TILE_STAGED")) s.tileStaged = std::strcmp(env, "0") != 0;
    if (!buildTileProgram(s, even, bestOwned, s.tileEven)) return;
    buildTileProgram(s, tilePasses(s, false, 0), bestOwned, s.tileOdd);
    s.useTiles = true;
}


// The reference lets a trajectory change only grip target/orientation/
// generation/selection mode and materials (tools/cloth_bag_metal.mm:4044-4090).
void checkTrajectoryConfig(const NumiClothHipSolver& s, const NumiClothBagGPUConfig& c) {
    const NumiClothBagGPUConfig& i = s.initialConfig;
    if (c.control.x != NUMI_CLOTH_BAG_GPU_ABI_VERSION || c.control.y != i.control.y || c.control.z != i.control.z ||
        c.control.w != i.control.w || c.gripControl.z != i.gripControl.z || c.gripControl.w != i.gripControl.w ||
        std::memcmp(&c.constraintCounts, &i.constraintCounts, sizeof(mr_uint4)) != 0 ||
        std::memcmp(&c.contactCounts, &i.contactCounts, sizeof(mr_uint4)) != 0 ||
        std::memcmp(&c.mouthControl, &i.mouthControl, sizeof(mr_uint4)) != 0) {
        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "cloth trajectory changes fixed topology");
    }
    if (c.gripControl.y == 1u && s.cuff.size() != 2ull * c.gripControl.z) {
        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "cloth trajectory has incomplete cuff topology");
    }
    if (c.gravityAndTimestep.w != i.gravityAndTimestep.w && s.extension.damping > 0.0f) {
        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "a damped scene keeps its timestep");
    }
    const bool air = c.airVelocityAndDensity.w > 0.0f &&
                     (c.aerodynamicCoefficients.x > 0.0f || c.aerodynamicCoefficients.y > 0.0f) && s.D != 0;
    if (air && !s.yarnAir) {
        fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "yarn air loads must be enabled by the scene's initial configuration");
    }
}

void stepSolver(NumiClothHipSolver& s, const NumiClothBagGPUConfig* configs, uint32_t substeps,
                uint32_t iterations, uint32_t sweeps) {
    if (substeps == 0) return;
    if (configs == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null configs");
    for (uint32_t k = 0; k < substeps; ++k) checkTrajectoryConfig(s, configs[k]);
    s.ensureStaging(substeps);
    // The pinned staging buffers are reused: the previous upload must be done.
    HIP_OK(hipStreamSynchronize(s.stream));
    std::memcpy(s.stagingConfigs, configs, substeps * sizeof(NumiClothBagGPUConfig));
    double t = s.time;
    for (uint32_t k = 0; k < substeps; ++k) {
        t += configs[k].gravityAndTimestep.w;
        s.stagingTimes[k] = t;
    }
    s.time = t;
    s.configs.upload(s.stagingConfigs, substeps, s.stream);
    s.times.upload(s.stagingTimes, substeps, s.stream);
    s.buildGraph(iterations, sweeps);
    const uint32_t zero = 0;
    (void)zero;
    k_reset_cursor<<<1, 64, 0, s.stream>>>(s.slot.ptr);
    // Status records count per step() call, like the reference's per-call buffers.
    k_fill_u32<<<1, 64, 0, s.stream>>>(reinterpret_cast<uint32_t*>(s.frictionStatus.ptr), 8, 0u);
    k_fill_u32<<<1, 64, 0, s.stream>>>(reinterpret_cast<uint32_t*>(s.selfStatus.ptr), 4, 0u);
    k_fill_u32<<<1, 64, 0, s.stream>>>(reinterpret_cast<uint32_t*>(s.aeroStatus.ptr), 4, 0u);
    HIP_OK(hipGetLastError());
    // With certificate passes the call's last substep runs outside the graph:
    // its certificate loop reads residuals back between rounds.
    const bool certify = s.extension.certificatePasses != 0;
    const uint32_t captured = certify ? substeps - 1 : substeps;
    s.certificateLast = 0;
    uint32_t done = 0;
    if (s.useGraphs && !s.launchDecided && captured >= kLaunchProbeSubsteps) {
        // One replay uploads the graph; then two substeps each way, timed.
        hipEvent_t marks[3];
        for (hipEvent_t& e : marks) HIP_OK(hipEventCreate(&e));
        HIP_OK(hipGraphLaunch(s.graph, s.stream));
        HIP_OK(hipEventRecord(marks[0], s.stream));
        for (int k = 0; k < 2; ++k) HIP_OK(hipGraphLaunch(s.graph, s.stream));
        HIP_OK(hipEventRecord(marks[1], s.stream));
        for (int k = 0; k < 2; ++k) s.recordSubstep(iterations, sweeps);
        HIP_OK(hipEventRecord(marks[2], s.stream));
        HIP_OK(hipEventSynchronize(marks[2]));
        HIP_OK(hipEventElapsedTime(&s.launchProbeGraphMs, marks[0], marks[1]));
        HIP_OK(hipEventElapsedTime(&s.launchProbeDirectMs, marks[1], marks[2]));
        for (hipEvent_t& e : marks) (void)hipEventDestroy(e);
        s.useGraphs = s.launchProbeGraphMs <= s.launchProbeDirectMs;
        s.launchDecided = true;
        done = kLaunchProbeSubsteps;
    }
    if (s.useGraphs) {
        for (uint32_t k = done; k < captured; ++k) HIP_OK(hipGraphLaunch(s.graph, s.stream));
    } else {
        for (uint32_t k = done; k < captured; ++k) s.recordSubstep(iterations, sweeps);
    }
    if (certify) s.recordSubstep(iterations, sweeps, nullptr, nullptr, true);
    HIP_OK(hipGetLastError());
}

// ------------------------------------------------- no-allocation step ----
// numi_cloth_hip_step_noalloc stages each call's substep configs and end
// times in `scratch`, device memory the library never allocates: a
// file-scope array that every device the library runs on holds until the
// process ends. Every call on a device shares that device's copy, so calls
// that overlap on one device race.
constexpr uint32_t kNoAllocSubsteps = 1024;
constexpr size_t kNoAllocConfigBytes = kNoAllocSubsteps * sizeof(NumiClothBagGPUConfig);
constexpr size_t kNoAllocScratchBytes = kNoAllocConfigBytes + kNoAllocSubsteps * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The device address of `scratch` on `device` (current), looked up on the
// first call there and kept: in host code the array's name is not a device
// address.
unsigned char* scratchAddress(int device) {
    static std::mutex lock;
    static std::array<unsigned char*, 64> addresses{};
    if (device < 0 || device >= static_cast<int>(addresses.size())) {
        fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "numi_cloth_hip_step_noalloc supports device indices 0-63");
    }
    std::lock_guard<std::mutex> hold(lock);
    if (addresses[device] == nullptr) {
        void* address = nullptr;
        const hipError_t status = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
        if (status != hipSuccess || address == nullptr) {
            fail(NUMI_CLOTH_HIP_ERROR_DEVICE,
                 std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(status));
        }
        addresses[device] = static_cast<unsigned char*>(address);
    }
    return addresses[device];
}

// Makes `device` current for a scope and restores the caller's device.
struct CurrentDevice {
    int previous = 0;
    bool changed = false;
    explicit CurrentDevice(int device) {
        HIP_OK(hipGetDevice(&previous));
        if (previous != device) {
            HIP_OK(hipSetDevice(device));
            changed = true;
        }
    }
    ~CurrentDevice() {
        if (changed) (void)hipSetDevice(previous);
    }
};

// step() without any allocation: at most kNoAllocSubsteps substeps per
// call, staged in `scratch`, launched directly (no graph capture, which
// allocates); the strain-limit program for `sweeps` must already exist (the
// solver prepares the reference's 3 sweeps at create, step() any other).
void stepSolverNoAlloc(NumiClothHipSolver& s, const NumiClothBagGPUConfig* configs, uint32_t substeps,
                       uint32_t iterations, uint32_t sweeps) {
    if (substeps == 0) return;
    if (configs == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null configs");
    if (substeps > kNoAllocSubsteps) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT,
             "numi_cloth_hip_step_noalloc takes at most " + std::to_string(kNoAllocSubsteps) +
                 " substeps per call, got " + std::to_string(substeps));
    }
    if (s.useTiles && sweeps != s.tileStrainSweeps) {
        fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED,
             "numi_cloth_hip_step_noalloc: the solver's strain-limit program is prepared for " +
                 std::to_string(s.tileStrainSweeps) + " sweeps, not " + std::to_string(sweeps) +
                 "; one numi_cloth_hip_step call with these sweeps prepares it");
    }
    for (uint32_t k = 0; k < substeps; ++k) checkTrajectoryConfig(s, configs[k]);
    CurrentDevice current(s.device);
    unsigned char* base = scratchAddress(s.device);
    auto* stagedConfigs = reinterpret_cast<NumiClothBagGPUConfig*>(base);
    auto* stagedTimes = reinterpret_cast<double*>(base + kNoAllocConfigBytes);
    std::array<double, kNoAllocSubsteps> times;
    double t = s.time;
    for (uint32_t k = 0; k < substeps; ++k) {
        t += configs[k].gravityAndTimestep.w;
        times[k] = t;
    }
    HIP_OK(hipMemcpyAsync(stagedConfigs, configs, substeps * sizeof(NumiClothBagGPUConfig), hipMemcpyHostToDevice,
                          s.stream));
    HIP_OK(hipMemcpyAsync(stagedTimes, times.data(), substeps * sizeof(double), hipMemcpyHostToDevice, s.stream));
    // The host arrays are the caller's and this frame's: the copies finish
    // before returning.
    HIP_OK(hipStreamSynchronize(s.stream));
    s.time = t;
    k_reset_cursor<<<1, 64, 0, s.stream>>>(s.slot.ptr);
    k_fill_u32<<<1, 64, 0, s.stream>>>(reinterpret_cast<uint32_t*>(s.frictionStatus.ptr), 8, 0u);
    k_fill_u32<<<1, 64, 0, s.stream>>>(reinterpret_cast<uint32_t*>(s.selfStatus.ptr), 4, 0u);
    k_fill_u32<<<1, 64, 0, s.stream>>>(reinterpret_cast<uint32_t*>(s.aeroStatus.ptr), 4, 0u);
    s.certificateLast = 0;
    for (uint32_t k = 0; k < substeps; ++k)
        s.recordSubstep(iterations, sweeps, stagedConfigs, stagedTimes,
                        k + 1 == substeps && s.extension.certificatePasses != 0);
    HIP_OK(hipGetLastError());
}

void readSolver(NumiClothHipSolver& s, const NumiClothHipState& out) {
    hipStream_t st = s.stream;
    std::vector<float4> hp, hq, hv;
    if (out.particles != nullptr) {
        hp.resize(s.P);
        hq.resize(s.P);
        hv.resize(s.P);
        s.pos.download(hp.data(), s.P, st);
        s.prev.download(hq.data(), s.P, st);
        s.vel.download(hv.data(), s.P, st);
    }
    std::vector<float> dl, bl, kl;
    if (out.distances != nullptr) {
        dl.resize(s.D);
        s.dLambda.download(dl.data(), s.D, st);
    }
    if (out.bends != nullptr) {
        bl.resize(s.B);
        s.bLambda.download(bl.data(), s.B, st);
    }
    if (out.knots != nullptr) {
        kl.resize(s.K);
        s.kLambda.download(kl.data(), s.K, st);
    }
    if (out.grips != nullptr) s.grips.download(out.grips, s.G, st);
    if (out.fruits != nullptr) s.fruits.download(out.fruits, s.F, st);
    if (out.fruitPairs != nullptr) s.fruitPairs.download(out.fruitPairs, s.FP, st);
    if (out.yarnContacts != nullptr) s.yarnContacts.download(out.yarnContacts, s.Y, st);
    if (out.selfStatus != nullptr) s.selfStatus.download(out.selfStatus, 1, st);
    if (out.frictionStatus != nullptr) s.frictionStatus.download(out.frictionStatus, 1, st);
    if (out.aerodynamicsStatus != nullptr) s.aeroStatus.download(out.aerodynamicsStatus, 1, st);
    if (out.releaseStatus != nullptr) s.releaseStatus.download(out.releaseStatus, 1, st);
    if (out.failure != nullptr) s.failure.download(out.failure, 1, st);
    std::vector<uint32_t> clearance;
    if (out.mouthClearance != nullptr && s.F != 0) {
        clearance.resize(s.F);
        s.mouthClearance.download(clearance.data(), s.F, st);
    }
    HIP_OK(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < clearance.size(); ++i) std::memcpy(out.mouthClearance + i, &clearance[i], sizeof(float));
    if (out.particles != nullptr) {
        for (uint32_t i = 0; i < s.P; ++i) {
            NumiClothBagGPUParticle& p = out.particles[i];
            p.positionAndInverseMass = {hp[i].x, hp[i].y, hp[i].z, hp[i].w};
            p.previousAndMass = {hq[i].x, hq[i].y, hq[i].z, hq[i].w};
            p.velocity = {hv[i].x, hv[i].y, hv[i].z, hv[i].w};
        }
    }
    if (out.distances != nullptr) {
        for (uint32_t k = 0; k < s.D; ++k) {
            out.distances[k] = s.hostDistances[k];
            out.distances[k].material.z = dl[k];
        }
    }
    if (out.bends != nullptr) {
        for (uint32_t k = 0; k < s.B; ++k) {
            out.bends[k] = s.hostBends[k];
            out.bends[k].material.w = bl[k];
        }
    }
    if (out.knots != nullptr) {
        for (uint32_t k = 0; k < s.K; ++k) {
            out.knots[k] = s.hostKnots[k];
            out.knots[k].material.z = kl[k];
        }
    }
}

void diagnoseSolver(NumiClothHipSolver& s, NumiClothHipDiagnostics* out) {
    const uint32_t n = std::max(s.P, s.D);
    const uint32_t blocks = std::min<uint32_t>(1024, blocksFor(n));
    const float radius = s.initialConfig.clothMaterial.x;
    k_diag_partial<<<blocks, kBlock, 0, s.stream>>>(s.pos.ptr, s.P, s.dIJ.ptr, s.dMaxLength.ptr, s.D, s.spheres.ptr,
                                                    s.extension.staticSphereCount, radius, s.diagPartials.ptr);
    if (s.self) {
        k_self_penetration_into<<<blocks, kBlock, 0, s.stream>>>(s.grid.ptr, s.pairs.ptr, s.dIJ.ptr, s.pos.ptr, radius,
                                                                 s.diagPartials.ptr);
    }
    k_diag_final<<<1, kBlock, 0, s.stream>>>(s.diagPartials.ptr, blocks, s.diagResult.ptr);
    HIP_OK(hipGetLastError());
    float4 r{};
    s.diagResult.download(&r, 1, s.stream);
    HIP_OK(hipStreamSynchronize(s.stream));
    std::memset(out, 0, sizeof(*out));
    out->maximumStrainViolation = r.x;
    out->maximumGroundPenetration = s.initialConfig.constraintCounts.z != 0u ? r.y : 0.0f;
    out->maximumSpherePenetration = r.z;
    out->maximumSelfPenetration = r.w;
    out->certificatePasses = s.certificateLast;
    {
        unsigned long long stats[2] = {0ull, 0ull};
        s.roundStats.download(stats, 2, s.stream);
        HIP_OK(hipStreamSynchronize(s.stream));
        out->tileRoundsSkipped = stats[0];
        out->tileRoundsRun = stats[1];
    }
    if (s.yarnStats.count != 0) {
        uint32_t stats[4] = {0u, 0u, 0u, 0u};
        s.yarnStats.download(stats, 4, s.stream);
        HIP_OK(hipStreamSynchronize(s.stream));
        out->yarnReruns = stats[0];
        out->yarnCandidatesMax = stats[1];
        out->yarnListBuilds = stats[2];
        out->yarnListSolves = stats[3];
    }
    if (s.self) {
        SelfGrid g{};
        s.grid.download(&g, 1, s.stream);
        HIP_OK(hipStreamSynchronize(s.stream));
        out->selfCandidateCount = g.valid ? g.pairCount : 0u;
        out->selfCandidateOverflow = g.pairTotal > g.pairCount || g.entryCount > g.sortCount ? 1u : 0u;
        out->selfEntryCount = g.valid ? g.entryCount : 0u;
        out->selfGridCells[0] = g.nx;
        out->selfGridCells[1] = g.ny;
        out->selfGridCells[2] = g.nz;
    }
    if (s.pointSelf) {
        s.pointContacts.download(&out->pointSelfContacts, 1, s.stream);
        HIP_OK(hipStreamSynchronize(s.stream));
    }
}

void timePass(NumiClothHipSolver& s, uint32_t pass, uint32_t colour, uint32_t repeats,
              NumiClothHipPassTiming* out) {
    if (repeats == 0) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "repeats must be positive");
    const std::vector<Range>* colours = nullptr;
    // Algorithmic bytes per item: distance 84 (ij, rest, compliance 20; lambda
    // read+write 8; two endpoints read 32, written 24), bend 80, strain check 48,
    // collide 28 (read xyzw 16, write xyz 12).
    double perItem = 0.0;
    if (pass == NUMI_CLOTH_HIP_PASS_DISTANCE) {
        colours = &s.distanceColours;
        perItem = 84.0;
    } else if (pass == NUMI_CLOTH_HIP_PASS_BEND) {
        colours = &s.bendColours;
        perItem = 80.0;
    } else if (pass == NUMI_CLOTH_HIP_PASS_STRAIN) {
        colours = &s.distanceColours;
        perItem = 48.0;
    } else if (pass != NUMI_CLOTH_HIP_PASS_COLLIDE && pass != NUMI_CLOTH_HIP_PASS_TILE_ITERATION) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "unknown pass");
    }
    if (pass == NUMI_CLOTH_HIP_PASS_TILE_ITERATION) {
        if (!s.useTiles || !s.tileEven.valid) fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "no tile program on this solver");
        out->averageMilliseconds = timeTileProgram(s, s.tileEven, repeats);
        out->algorithmicBytes = s.tileEven.algorithmicBytes;
        out->launches = repeats;
        out->itemsPerLaunch = static_cast<uint32_t>(s.tileEven.unique);
        return;
    }
    Range range{0, s.P};
    if (colours != nullptr) {
        if (colour >= colours->size()) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "colour out of range");
        range = (*colours)[colour];
    } else {
        perItem = 28.0;
    }
    hipEvent_t start = nullptr, stop = nullptr;
    HIP_OK(hipEventCreate(&start));
    HIP_OK(hipEventCreate(&stop));
    HIP_OK(hipEventRecord(start, s.stream));
    for (uint32_t k = 0; k < repeats; ++k) {
        if (pass == NUMI_CLOTH_HIP_PASS_DISTANCE) {
            k_distance<<<blocksFor(range.count), kBlock, 0, s.stream>>>(s.cfg(), s.ok(), range.first, range.count,
                                                                        s.pos.ptr, s.dIJ.ptr, s.dRestCompliance.ptr,
                                                                        s.dLambda.ptr, s.failure.ptr);
        } else if (pass == NUMI_CLOTH_HIP_PASS_BEND) {
            k_bend<<<blocksFor(range.count), kBlock, 0, s.stream>>>(s.cfg(), s.ok(), range.first, range.count,
                                                                    s.pos.ptr, s.bEnds.ptr, s.bChordCompliance.ptr,
                                                                    s.bLambda.ptr, s.failure.ptr);
        } else if (pass == NUMI_CLOTH_HIP_PASS_STRAIN) {
            k_strain<<<blocksFor(range.count), kBlock, 0, s.stream>>>(s.ok(), range.first, range.count, s.pos.ptr,
                                                                      s.dIJ.ptr, s.dMaxLength.ptr, s.failure.ptr);
        } else {
            const uint32_t n = s.extension.staticSphereCount;
            k_collide<<<blocksFor(s.P), kBlock, 2 * n * sizeof(float4), s.stream>>>(s.cfg(), s.ok(), s.pos.ptr, s.P,
                                                                               s.spheres.ptr, n, s.sphereWindowMin,
                                                                               s.failure.ptr);
        }
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(stop, s.stream));
    HIP_OK(hipEventSynchronize(stop));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    out->averageMilliseconds = static_cast<double>(ms) / repeats;
    out->algorithmicBytes = perItem * range.count;
    out->launches = repeats;
    out->itemsPerLaunch = range.count;
}

void selfBroadPhase(NumiClothHipSolver& s, NumiClothHipSelfBroadPhase* io) {
    if (!s.self) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "the scene has no hash self-contact");
    SubstepSlot slot{};
    slot.config = s.initialConfig;
    slot.ok = 1u;
    s.debugSlot.upload(&slot, 1, s.stream);
    s.launchBroadPhase(&s.debugSlot.ptr->config, &s.debugSlot.ptr->ok);
    HIP_OK(hipGetLastError());
    SelfGrid g{};
    s.grid.download(&g, 1, s.stream);
    HIP_OK(hipStreamSynchronize(s.stream));
    io->pairCount = g.valid ? g.pairCount : 0u;
    io->entryCount = g.valid ? g.sortCount : 0u;
    io->origin[0] = g.origin.x;
    io->origin[1] = g.origin.y;
    io->origin[2] = g.origin.z;
    io->cell = g.cell;
    io->cells[0] = g.nx;
    io->cells[1] = g.ny;
    io->cells[2] = g.nz;
    io->valid = g.valid;
    if (io->pairs != nullptr) {
        const uint32_t n = std::min(io->pairCount, io->pairCapacity);
        s.pairs.download(reinterpret_cast<uint2*>(io->pairs), n, s.stream);
    }
    // Both sorts leave the sorted entries in A (the counting sort's candidate
    // pass, or its rank pass, writes A; the radix sort's four ping-pong passes
    // end there).
    if (io->cellIds != nullptr) s.keysA.download(io->cellIds, std::min(io->entryCount, io->entryCapacity), s.stream);
    if (io->segments != nullptr) s.valsA.download(io->segments, std::min(io->entryCount, io->entryCapacity), s.stream);
    HIP_OK(hipStreamSynchronize(s.stream));
    if (io->segments != nullptr)  // drop the owner-cell flags kept above the segment index
        for (uint32_t k = 0; k < std::min(io->entryCount, io->entryCapacity); ++k) io->segments[k] &= kEntrySegmentMask;
}

}  // namespace

void NumiClothHipSolver::ensureStrainProgram(uint32_t sweeps) {
    if (!useTiles || sweeps == tileStrainSweeps) return;
    tileStrainSweeps = sweeps;
    if (sweeps == 0 || !buildTileProgram(*this, tilePasses(*this, false, sweeps), tileOwned, tileStrain)) {
        tileStrain.valid = false;
    } else {
        // Every tile of every round writes its flag, so they need no clearing.
        roundFlags.allocate(static_cast<size_t>(kReconciliationPasses + kCertificatePasses + 1) * tileStrain.tiles);
    }
    if (graph != nullptr) {
        (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
}

// ------------------------------------------------------------------ batch --
// Many independent specimens stepped by one solver over their merged scene:
// every colour launch, tile program and hipGraph replay covers all specimens.

// Device buffers of the specimen adjoint (created on first use).
struct AdjointBuffers {
    uint32_t distanceColours = 0, bendColours = 0, maxKnots = 0, maxConstraints = 0;
    size_t ldsBytes = 0;
    DeviceArray<SpecimenInfo> info;
    DeviceArray<uint4> distanceRanges, bendRanges;
    DeviceArray<NumiClothBagGPUConfig> configs;
    DeviceArray<double> times;
    DeviceArray<uint32_t> tracked;
    DeviceArray<float> targets;
    DeviceArray<float4> checkpointPos, checkpointVel;
    // One segment's recomputation: scratch state and the tape it records.
    // Two sets, so that the recomputation of segment k - 1 (side stream) runs
    // while segment k is reversed (solver stream).
    struct TapeSet {
        DeviceArray<float4> scratchPos, scratchPrev, scratchVel;
        DeviceArray<float> scratchLambdaD, scratchLambdaB;
        DeviceArray<float4> tapeX, tapeV0;
        DeviceArray<float> tapeLambdaD, tapeLambdaB, lossGrad;
        DeviceArray<uint32_t> tapeActive;
        DeviceArray<uint8_t> tapeStrainActs;
    };
    TapeSet sets[2];
    hipStream_t side = nullptr;                 // default priority, like the solver stream
    hipEvent_t recorded[2] = {nullptr, nullptr};  // set k's tape is written
    hipEvent_t released[2] = {nullptr, nullptr};  // set k's tape has been reversed
    DeviceArray<float4> gxState, gvState;
    DeviceArray<float> gradInverseMass, gradDamping, gradDistance, gradBend;
    DeviceArray<double> loss;
    DeviceArray<uint32_t> strainTaped;          // strain passes taped by the last adjoint call
    uint32_t lastSubsteps = 0, lastPasses = 0;  // of the last adjoint call
    std::vector<uint32_t> trackedFirst, trackedCount;
    std::vector<SpecimenInfo> baseInfo;         // per specimen, no tracked knots
    std::vector<NumiClothBagGPUConfig> hostConfigs;
    std::vector<double> hostTimes;
    bool specimensReady = false;                // layout for the specimen kernels
    bool ready = false;                         // adjoint buffers
    ~AdjointBuffers() {
        for (int k = 0; k < 2; ++k) {
            if (recorded[k] != nullptr) (void)hipEventDestroy(recorded[k]);
            if (released[k] != nullptr) (void)hipEventDestroy(released[k]);
        }
        if (side != nullptr) (void)hipStreamDestroy(side);
    }
};

struct NumiClothHipBatch {
    NumiClothHipScene* merged = nullptr;
    std::unique_ptr<NumiClothHipSolver> solver;
    uint32_t specimens = 0;
    std::vector<NumiClothBagGPUConfig> configs;
    uint64_t version = 0;          // step() calls so far
    uint64_t lambdaVersion = ~0ull; // version of the cached multipliers
    std::vector<float> dl, bl, kl;
    AdjointBuffers adjoint;
    ~NumiClothHipBatch() {
        solver.reset();
        if (merged != nullptr) numi_cloth_scene_destroy(merged);
    }
};

namespace {

void stepBatch(NumiClothHipBatch& b, const NumiClothBagGPUConfig* configs, uint32_t substeps, uint32_t iterations,
               uint32_t sweeps) {
    if (substeps == 0) return;
    if (configs == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null configs");
    const NumiClothBagGPUConfig& totals = b.merged->config;
    b.configs.assign(configs, configs + substeps);
    for (NumiClothBagGPUConfig& c : b.configs) {
        c.control.y = totals.control.y;
        c.control.z = totals.control.z;
        c.control.w = totals.control.w;
        c.constraintCounts = totals.constraintCounts;
        c.contactCounts = totals.contactCounts;
    }
    stepSolver(*b.solver, b.configs.data(), substeps, iterations, sweeps);
    ++b.version;
}

const uint32_t* mergeIndices(const NumiClothHipBatch& b, uint32_t family, uint32_t specimen, uint32_t* count) {
    const uint32_t* indices = nullptr;
    if (numi_cloth_scene_merge_indices(b.merged, family, specimen, &indices, count) != 0) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "specimen out of range");
    }
    return indices;
}

void readBatch(NumiClothHipBatch& b, uint32_t specimen, const NumiClothHipState& out) {
    if (specimen >= b.specimens) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "specimen out of range");
    NumiClothHipSolver& s = *b.solver;
    hipStream_t st = s.stream;
    uint32_t particles = 0;
    const uint32_t* particleIndex = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_PARTICLE, specimen, &particles);
    const uint32_t first = particles != 0 ? particleIndex[0] : 0u;
    std::vector<float4> hp(particles), hq(particles), hv(particles);
    if (out.particles != nullptr) {
        s.pos.downloadRange(hp.data(), first, particles, st);
        s.prev.downloadRange(hq.data(), first, particles, st);
        s.vel.downloadRange(hv.data(), first, particles, st);
    }
    // Multipliers are gathered from full downloads, cached per step() call.
    const bool needLambda = out.distances != nullptr || out.bends != nullptr || out.knots != nullptr;
    if (needLambda && b.lambdaVersion != b.version) {
        b.dl.resize(s.D);
        b.bl.resize(s.B);
        b.kl.resize(s.K);
        s.dLambda.download(b.dl.data(), s.D, st);
        s.bLambda.download(b.bl.data(), s.B, st);
        s.kLambda.download(b.kl.data(), s.K, st);
        b.lambdaVersion = b.version;
    }
    if (out.failure != nullptr) s.failure.download(out.failure, 1, st);
    if (out.frictionStatus != nullptr) s.frictionStatus.download(out.frictionStatus, 1, st);
    HIP_OK(hipStreamSynchronize(st));
    if (out.particles != nullptr) {
        for (uint32_t i = 0; i < particles; ++i) {
            NumiClothBagGPUParticle& p = out.particles[i];
            p.positionAndInverseMass = {hp[i].x, hp[i].y, hp[i].z, hp[i].w};
            p.previousAndMass = {hq[i].x, hq[i].y, hq[i].z, hq[i].w};
            p.velocity = {hv[i].x, hv[i].y, hv[i].z, hv[i].w};
        }
    }
    uint32_t count = 0;
    if (out.distances != nullptr) {
        const uint32_t* index = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_DISTANCE, specimen, &count);
        for (uint32_t k = 0; k < count; ++k) {
            NumiClothBagGPUDistance d = b.merged->distances[index[k]];
            d.particlesAndColor.x -= first;
            d.particlesAndColor.y -= first;
            d.material.z = b.dl[index[k]];
            out.distances[k] = d;
        }
    }
    if (out.bends != nullptr) {
        const uint32_t* index = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_BEND, specimen, &count);
        for (uint32_t k = 0; k < count; ++k) {
            NumiClothBagGPUBend d = b.merged->bends[index[k]];
            d.particlesAndColor.x -= first;
            d.particlesAndColor.y -= first;
            d.particlesAndColor.z -= first;
            d.material.w = b.bl[index[k]];
            out.bends[k] = d;
        }
    }
    if (out.knots != nullptr) {
        const uint32_t* index = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_KNOT, specimen, &count);
        for (uint32_t k = 0; k < count; ++k) {
            NumiClothBagGPUKnot d = b.merged->knots[index[k]];
            d.particles.x -= first;
            d.particles.y -= first;
            d.particles.z -= first;
            d.particles.w -= first;
            d.material.z = b.kl[index[k]];
            out.knots[k] = d;
        }
    }
}

// ---------------------------------------------------------------- adjoint --

// Per-specimen colour ranges of the merged scene (colour c of specimen s is
// one contiguous range of merged colour batch c) and the LDS footprint of the
// specimen kernels (one workgroup per specimen, state resident in LDS).
void setupSpecimens(NumiClothHipBatch& b) {
    AdjointBuffers& A = b.adjoint;
    if (A.specimensReady) return;
    NumiClothHipSolver& s = *b.solver;
    const NumiClothHipSceneDesc& d = b.merged->desc;
    const NumiClothBagGPUConfig& c = *d.config;
    if (d.knotCount || d.gripCount || d.fruitCount || d.extension.staticSphereCount || d.extension.selfContact ||
        c.constraintCounts.z != 0u || c.airVelocityAndDensity.w > 0.0f) {
        fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED,
             "the specimen kernels cover distances, bends, strain limits, drive and damping (no contacts, grips, "
             "crossing knots, fruits or air)");
    }
    const uint32_t S = b.specimens;
    std::vector<uint32_t> owner(s.P);
    std::vector<SpecimenInfo> info(S);
    for (uint32_t k = 0; k < S; ++k) {
        uint32_t count = 0;
        const uint32_t* idx = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_PARTICLE, k, &count);
        info[k] = SpecimenInfo{count ? idx[0] : 0u, count, 0u, 0u, 0u, 0u, {0u, 0u}};
        for (uint32_t i = 0; i < count; ++i) owner[idx[i]] = k;
        if (count > kSpecimenOwn * kSpecimenThreads) {
            fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "specimens hold at most 5120 knots");
        }
    }
    auto ranges = [&](const NumiClothBagGPUBatch* batches, uint32_t batchCount, auto endsOf, uint32_t* colours,
                      bool distance) {
        uint32_t maxColour = 0;
        for (uint32_t k = 0; k < batchCount; ++k) maxColour = std::max(maxColour, batches[k].control.z + 1);
        *colours = maxColour;
        std::vector<uint4> out(static_cast<size_t>(S) * maxColour, make_uint4(0u, 0u, 0u, 0u));
        std::vector<uint32_t> local(S, 0u);
        for (uint32_t k = 0; k < batchCount; ++k) {
            const mr_uint4 batch = batches[k].control;
            uint32_t j = batch.x;
            while (j < batch.x + batch.y) {
                const uint32_t spec = owner[endsOf(j)];
                uint32_t end = j;
                while (end < batch.x + batch.y && owner[endsOf(end)] == spec) ++end;
                uint4& r = out[static_cast<size_t>(spec) * maxColour + batch.z];
                if (r.y != 0) fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, "specimen colour split across a batch");
                r = make_uint4(j, end - j, local[spec], 0u);
                local[spec] += end - j;
                j = end;
            }
        }
        for (uint32_t k = 0; k < S; ++k) (distance ? info[k].distanceCount : info[k].bendCount) = local[k];
        return out;
    };
    const auto dr = ranges(d.distanceBatches, d.distanceBatchCount,
                           [&](uint32_t j) { return d.distances[j].particlesAndColor.x; }, &A.distanceColours, true);
    const auto br = ranges(d.bendBatches, d.bendBatchCount,
                           [&](uint32_t j) { return d.bends[j].particlesAndColor.x; }, &A.bendColours, false);
    A.maxKnots = 0;
    A.maxConstraints = 0;
    for (const SpecimenInfo& i : info) {
        A.maxKnots = std::max(A.maxKnots, i.knotCount);
        A.maxConstraints = std::max(A.maxConstraints, i.distanceCount + i.bendCount);
    }
    int maxShared = 0;
    HIP_OK(hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, s.device));
    A.ldsBytes = static_cast<size_t>(A.maxKnots) * sizeof(float4) + static_cast<size_t>(A.maxConstraints) * 4;
    if (A.ldsBytes + 1024 > static_cast<size_t>(maxShared)) {
        fail(NUMI_CLOTH_HIP_ERROR_UNSUPPORTED, "specimen does not fit in one workgroup's LDS");
    }
    hipStream_t st = s.stream;
    A.baseInfo = info;
    A.info.allocate(S);
    A.info.upload(A.baseInfo.data(), S, st);
    A.distanceRanges.allocate(dr.size());
    A.distanceRanges.upload(dr.data(), dr.size(), st);
    A.bendRanges.allocate(std::max<size_t>(br.size(), 1));
    A.bendRanges.upload(br.data(), br.size(), st);
    A.loss.allocate(S);
    A.tracked.allocate(1);
    A.targets.allocate(1);
    HIP_OK(hipStreamSynchronize(st));
    A.specimensReady = true;
}

// Gradient accumulators, tape sets and the side stream of the adjoint.
void setupAdjoint(NumiClothHipBatch& b) {
    setupSpecimens(b);
    AdjointBuffers& A = b.adjoint;
    if (A.ready) return;
    NumiClothHipSolver& s = *b.solver;
    hipStream_t st = s.stream;
    A.gxState.allocate(s.P);
    A.gvState.allocate(s.P);
    A.gradInverseMass.allocate(s.P);
    A.gradDamping.allocate(s.P);
    A.gradDistance.allocate(std::max(s.D, 1u));
    A.gradBend.allocate(std::max(s.B, 1u));
    A.strainTaped.allocate(1);
    for (AdjointBuffers::TapeSet& set : A.sets) {
        set.scratchPos.allocate(s.P);
        set.scratchPrev.allocate(s.P);
        set.scratchVel.allocate(s.P);
        set.scratchLambdaD.allocate(std::max(s.D, 1u));
        set.scratchLambdaB.allocate(std::max(s.B, 1u));
    }
    HIP_OK(hipSetDevice(s.device));
    HIP_OK(hipStreamCreateWithFlags(&A.side, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        HIP_OK(hipEventCreateWithFlags(&A.recorded[k], hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&A.released[k], hipEventDisableTiming));
    }
    HIP_OK(hipStreamSynchronize(st));
    A.ready = true;
}

uint32_t adjointPasses(const AdjointBuffers& A, uint32_t iterations, uint32_t sweeps) {
    return A.distanceColours * iterations + A.bendColours * ((iterations + 1) / 2) +
           (sweeps == 0 ? 0u : 18u * sweeps * A.distanceColours);
}

// The trajectory's configs with the batch totals and the time at the end of
// every substep, uploaded for the specimen kernels; returns the end time.
double uploadSpecimenConfigs(NumiClothHipBatch& b, const NumiClothBagGPUConfig* in, uint32_t T) {
    NumiClothHipSolver& s = *b.solver;
    AdjointBuffers& A = b.adjoint;
    A.hostConfigs.assign(in, in + T);
    A.hostTimes.resize(T);
    double t = s.time;
    const NumiClothBagGPUConfig& totals = b.merged->config;
    for (uint32_t k = 0; k < T; ++k) {
        NumiClothBagGPUConfig& c = A.hostConfigs[k];
        c.control.y = totals.control.y;
        c.control.z = totals.control.z;
        c.control.w = totals.control.w;
        c.constraintCounts = totals.constraintCounts;
        c.contactCounts = totals.contactCounts;
        checkTrajectoryConfig(s, c);
        if (c.control.x != NUMI_CLOTH_BAG_GPU_ABI_VERSION || !(c.gravityAndTimestep.w > 0.0f) ||
            !std::isfinite(c.gravityAndTimestep.w) || !std::isfinite(c.gravityAndTimestep.z)) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "specimen configs must be valid (ABI version, finite dt > 0)");
        }
        t += c.gravityAndTimestep.w;
        A.hostTimes[k] = t;
    }
    if (A.configs.count < T) A.configs.allocate(T);
    A.configs.upload(A.hostConfigs.data(), T, s.stream);
    if (A.times.count < T) A.times.allocate(T);
    A.times.upload(A.hostTimes.data(), T, s.stream);
    return t;
}

// Fields of the specimen kernels' arguments that every launch shares; the
// state pointers are the solver's own.
SpecimenArgs specimenArgs(NumiClothHipBatch& b, uint32_t iterations, uint32_t sweeps) {
    NumiClothHipSolver& s = *b.solver;
    AdjointBuffers& A = b.adjoint;
    SpecimenArgs a{};
    a.specimens = A.info.ptr;
    a.distanceRanges = A.distanceRanges.ptr;
    a.bendRanges = A.bendRanges.ptr;
    a.distanceColours = A.distanceColours;
    a.bendColours = A.bendColours;
    a.configs = A.configs.ptr;
    a.times = A.times.ptr;
    a.iterations = iterations;
    a.sweeps = sweeps;
    a.anchor = s.driven ? s.anchor.ptr : nullptr;
    a.motion = s.motion.ptr;
    a.drive = make_float4(s.extension.driveAmplitude, s.extension.driveOmega, s.extension.drivePhase, 0.0f);
    const float dt0 = s.initialConfig.gravityAndTimestep.w;
    a.dampingFactor = s.extension.damping > 0.0f ? 1.0f / (1.0f + s.extension.damping * dt0) : 1.0f;
    a.dampingDt = dt0;
    a.ij = s.dIJ.ptr;
    a.restCompliance = s.dRestCompliance.ptr;
    a.maxLength = s.dMaxLength.ptr;
    a.bEnds = s.bEnds.ptr;
    a.chordCompliance = s.bChordCompliance.ptr;
    a.tracked = A.tracked.ptr;
    a.targets = A.targets.ptr;
    a.loss = A.loss.ptr;
    a.passes = adjointPasses(A, iterations, sweeps);
    a.particles = s.P;
    a.distances = s.D;
    a.bends = s.B;
    a.specimenCount = b.specimens;
    a.failure = s.failure.ptr;
    a.pos = s.pos.ptr;
    a.prev = s.prev.ptr;
    a.vel = s.vel.ptr;
    a.dLambda = s.dLambda.ptr;
    a.bLambda = s.bLambda.ptr;
    return a;
}

// One persistent launch for all substeps of the call: one workgroup per
// specimen, its knots and multipliers resident in LDS throughout (the
// adjoint's forward kernel with neither tape nor tracked knots). Leaves the
// batch where stepBatch would, bit for bit.
void stepResident(NumiClothHipBatch& b, const NumiClothBagGPUConfig* configs, uint32_t T, uint32_t iterations,
                  uint32_t sweeps) {
    if (T == 0) return;
    if (configs == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null configs");
    NumiClothHipSolver& s = *b.solver;
    setupSpecimens(b);
    AdjointBuffers& A = b.adjoint;
    const double end = uploadSpecimenConfigs(b, configs, T);
    A.info.upload(A.baseInfo.data(), b.specimens, s.stream);  // no tracked knots
    SpecimenArgs f = specimenArgs(b, iterations, sweeps);
    f.firstSubstep = 0;
    f.substeps = T;
    f.record = 0u;
    f.trackedTotal = 0;
    k_specimen_forward<<<b.specimens, kSpecimenThreads, A.ldsBytes, s.stream>>>(f);
    HIP_OK(hipGetLastError());
    s.time = end;
    ++b.version;
}

void runAdjoint(NumiClothHipBatch& b, const NumiClothHipAdjointProblem& pr) {
    NumiClothHipSolver& s = *b.solver;
    setupAdjoint(b);
    AdjointBuffers& A = b.adjoint;
    const uint32_t T = pr.substeps, L = pr.checkpointInterval, S = b.specimens;
    if (T == 0 || L == 0 || pr.configs == nullptr || (pr.trackedPerSpecimen != 0 && pr.tracked == nullptr) ||
        (pr.trackedPerSpecimen != 0 && pr.targets == nullptr)) {
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "adjoint problem needs substeps, a checkpoint interval, configs, "
                                                    "tracked knots and targets");
    }
    if (pr.observable != NUMI_CLOTH_ADJOINT_TRACKING && pr.observable != NUMI_CLOTH_ADJOINT_LINEAR)
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "adjoint observable must be TRACKING or LINEAR");
    hipStream_t st = s.stream;
    HIP_OK(hipStreamSynchronize(st));
    const double t = uploadSpecimenConfigs(b, pr.configs, T);
    const uint32_t Tk = pr.trackedPerSpecimen;
    std::vector<uint32_t> tracked(static_cast<size_t>(S) * Tk);
    std::vector<SpecimenInfo> info = A.baseInfo;
    for (uint32_t k = 0; k < S; ++k) {
        for (uint32_t q = 0; q < Tk; ++q) {
            const uint32_t local = pr.tracked[static_cast<size_t>(k) * Tk + q];
            if (local >= info[k].knotCount) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "tracked knot out of range");
            tracked[static_cast<size_t>(k) * Tk + q] = info[k].knotFirst + local;
        }
        info[k].trackedFirst = k * Tk;
        info[k].trackedCount = Tk;
    }
    A.info.upload(info.data(), S, st);
    // Buffers grow to the largest call and are reused by later calls.
    A.tracked.ensure(std::max<size_t>(tracked.size(), 1));
    A.tracked.upload(tracked.data(), tracked.size(), st);
    const size_t targetCount = static_cast<size_t>(T) * S * Tk * 3;
    A.targets.ensure(std::max<size_t>(targetCount, 1));
    A.targets.upload(pr.targets, targetCount, st);
    const uint32_t segments = (T + L - 1) / L;
    const uint32_t passes = adjointPasses(A, pr.iterations, pr.strainSweeps);
    A.checkpointPos.ensure(static_cast<size_t>(segments) * s.P);
    A.checkpointVel.ensure(static_cast<size_t>(segments) * s.P);
    A.lastSubsteps = T;
    A.lastPasses = pr.strainSweeps == 0 ? 0u : 18u * pr.strainSweeps * A.distanceColours;
    const uint32_t setsUsed = segments > 1 ? 2u : 1u;
    for (uint32_t k = 0; k < setsUsed; ++k) {
        AdjointBuffers::TapeSet& set = A.sets[k];
        set.tapeX.ensure(static_cast<size_t>(L) * passes * s.P);
        set.tapeV0.ensure(static_cast<size_t>(L) * s.P);
        set.tapeLambdaD.ensure(std::max<size_t>(static_cast<size_t>(L) * pr.iterations * s.D, 1));
        set.tapeLambdaB.ensure(std::max<size_t>(static_cast<size_t>(L) * ((pr.iterations + 1) / 2) * s.B, 1));
        set.tapeActive.ensure(std::max<size_t>(static_cast<size_t>(L) * passes * S, 1));
        set.tapeStrainActs.ensure(std::max<size_t>(static_cast<size_t>(L) * A.lastPasses * s.D, 1));
        set.lossGrad.ensure(std::max<size_t>(static_cast<size_t>(L) * S * Tk * 3, 1));
    }
    HIP_OK(hipMemsetAsync(A.loss.ptr, 0, S * sizeof(double), st));
    HIP_OK(hipMemsetAsync(A.gxState.ptr, 0, s.P * sizeof(float4), st));
    HIP_OK(hipMemsetAsync(A.gvState.ptr, 0, s.P * sizeof(float4), st));
    HIP_OK(hipMemsetAsync(A.gradInverseMass.ptr, 0, s.P * sizeof(float), st));
    HIP_OK(hipMemsetAsync(A.gradDamping.ptr, 0, s.P * sizeof(float), st));
    HIP_OK(hipMemsetAsync(A.gradDistance.ptr, 0, std::max(s.D, 1u) * sizeof(float), st));
    HIP_OK(hipMemsetAsync(A.gradBend.ptr, 0, std::max(s.B, 1u) * sizeof(float), st));
    HIP_OK(hipMemsetAsync(A.strainTaped.ptr, 0, sizeof(uint32_t), st));

    SpecimenArgs a = specimenArgs(b, pr.iterations, pr.strainSweeps);
    a.trackedTotal = S * Tk;
    a.linear = pr.observable == NUMI_CLOTH_ADJOINT_LINEAR ? 1u : 0u;
    a.gxState = A.gxState.ptr;
    a.gvState = A.gvState.ptr;
    a.gradInverseMass = A.gradInverseMass.ptr;
    a.gradDamping = A.gradDamping.ptr;
    a.gradDistance = A.gradDistance.ptr;
    a.gradBend = A.gradBend.ptr;
    a.strainTaped = A.strainTaped.ptr;

    // The tape arguments of set `which` for a recording forward.
    auto taped = [&](SpecimenArgs f, uint32_t which) {
        AdjointBuffers::TapeSet& set = A.sets[which];
        f.record = 1u;
        f.tapeX = set.tapeX.ptr;
        f.tapeLambdaD = set.tapeLambdaD.ptr;
        f.tapeLambdaB = set.tapeLambdaB.ptr;
        f.tapeV0 = set.tapeV0.ptr;
        f.tapeActive = set.tapeActive.ptr;
        f.tapeStrainActs = set.tapeStrainActs.ptr;
        f.lossGrad = set.lossGrad.ptr;
        return f;
    };
    // Forward over the trajectory on the solver's own state, a checkpoint of
    // (x, v) at the start of every segment. The last segment records its tape
    // as it goes (into the set its reverse reads first), so the reverse
    // starts without recomputing it.
    for (uint32_t k = 0; k < segments; ++k) {
        HIP_OK(hipMemcpyAsync(A.checkpointPos.ptr + static_cast<size_t>(k) * s.P, s.pos.ptr, s.P * sizeof(float4),
                              hipMemcpyDeviceToDevice, st));
        HIP_OK(hipMemcpyAsync(A.checkpointVel.ptr + static_cast<size_t>(k) * s.P, s.vel.ptr, s.P * sizeof(float4),
                              hipMemcpyDeviceToDevice, st));
        SpecimenArgs f = a;
        f.firstSubstep = k * L;
        f.substeps = std::min(L, T - k * L);
        f.record = 0u;
        f.pos = s.pos.ptr;
        f.prev = s.prev.ptr;
        f.vel = s.vel.ptr;
        f.dLambda = s.dLambda.ptr;
        f.bLambda = s.bLambda.ptr;
        if (k + 1 == segments) {
            f = taped(f, 0u);
            f.record = 2u;
        }
        k_specimen_forward<<<S, kSpecimenThreads, A.ldsBytes, st>>>(f);
        HIP_OK(hipGetLastError());
    }
    // Reverse, segment by segment: recompute the segment from its checkpoint
    // with the tape on (scratch state), then run it backwards. The
    // recomputations run on the side stream into alternating tape sets, one
    // segment ahead of the reverse on the solver stream: recomputing segment
    // k waits for the checkpoints and for the reverse that last read its set
    // (segment k + 2); reversing segment k waits for its recomputation.
    HIP_OK(hipEventRecord(A.released[0], st));  // checkpoints written
    HIP_OK(hipEventRecord(A.released[1], st));
    for (uint32_t k = segments; k-- > 0;) {
        const uint32_t which = (segments - 1 - k) % setsUsed;
        AdjointBuffers::TapeSet& set = A.sets[which];
        if (k + 1 == segments) {
            // Taped by the forward itself (set 0, on the solver stream).
            SpecimenArgs f = taped(a, which);
            f.firstSubstep = k * L;
            f.substeps = std::min(L, T - k * L);
            f.dLambda = s.dLambda.ptr;
            f.bLambda = s.bLambda.ptr;
            f.pos = s.pos.ptr;  // inverse masses for the reverse
            f.prev = s.prev.ptr;
            f.vel = s.vel.ptr;
            k_specimen_reverse<<<S, kSpecimenThreads, A.ldsBytes, st>>>(f);
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventRecord(A.released[which], st));
            continue;
        }
        hipStream_t rs = setsUsed > 1 ? A.side : st;
        HIP_OK(hipStreamWaitEvent(rs, A.released[which], 0));
        HIP_OK(hipMemcpyAsync(set.scratchPos.ptr, A.checkpointPos.ptr + static_cast<size_t>(k) * s.P,
                              s.P * sizeof(float4), hipMemcpyDeviceToDevice, rs));
        HIP_OK(hipMemcpyAsync(set.scratchVel.ptr, A.checkpointVel.ptr + static_cast<size_t>(k) * s.P,
                              s.P * sizeof(float4), hipMemcpyDeviceToDevice, rs));
        // prev is written before it is read; its w lane (the mass) is kept.
        HIP_OK(hipMemcpyAsync(set.scratchPrev.ptr, s.prev.ptr, s.P * sizeof(float4), hipMemcpyDeviceToDevice, rs));
        SpecimenArgs f = a;
        f.firstSubstep = k * L;
        f.substeps = std::min(L, T - k * L);
        f.record = 1u;
        f.pos = set.scratchPos.ptr;
        f.prev = set.scratchPrev.ptr;
        f.vel = set.scratchVel.ptr;
        f.dLambda = set.scratchLambdaD.ptr;
        f.bLambda = set.scratchLambdaB.ptr;
        f.tapeX = set.tapeX.ptr;
        f.tapeLambdaD = set.tapeLambdaD.ptr;
        f.tapeLambdaB = set.tapeLambdaB.ptr;
        f.tapeV0 = set.tapeV0.ptr;
        f.tapeActive = set.tapeActive.ptr;
        f.tapeStrainActs = set.tapeStrainActs.ptr;
        f.lossGrad = set.lossGrad.ptr;
        k_specimen_forward<<<S, kSpecimenThreads, A.ldsBytes, rs>>>(f);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(A.recorded[which], rs));
        HIP_OK(hipStreamWaitEvent(st, A.recorded[which], 0));
        f.pos = s.pos.ptr;  // inverse masses for the reverse
        k_specimen_reverse<<<S, kSpecimenThreads, A.ldsBytes, st>>>(f);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(A.released[which], st));
    }
    HIP_OK(hipStreamSynchronize(st));
    HIP_OK(hipStreamSynchronize(A.side));
    s.time = t;
    ++b.version;
}

void readGradients(NumiClothHipBatch& b, uint32_t specimen, const NumiClothHipSpecimenGradients& out) {
    if (!b.adjoint.ready) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "no adjoint has run on this batch");
    if (specimen >= b.specimens) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "specimen out of range");
    NumiClothHipSolver& s = *b.solver;
    AdjointBuffers& A = b.adjoint;
    hipStream_t st = s.stream;
    std::vector<float> gd(s.D), gb(s.B), gw(s.P), gc(s.P);
    std::vector<double> loss(b.specimens);
    A.gradDistance.download(gd.data(), s.D, st);
    A.gradBend.download(gb.data(), s.B, st);
    A.gradInverseMass.download(gw.data(), s.P, st);
    A.gradDamping.download(gc.data(), s.P, st);
    A.loss.download(loss.data(), b.specimens, st);
    HIP_OK(hipStreamSynchronize(st));
    uint32_t count = 0;
    if (out.loss != nullptr) *out.loss = loss[specimen];
    const uint32_t* index = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_DISTANCE, specimen, &count);
    if (out.distanceCompliance != nullptr)
        for (uint32_t k = 0; k < count; ++k) out.distanceCompliance[k] = gd[index[k]];
    index = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_BEND, specimen, &count);
    if (out.bendCompliance != nullptr)
        for (uint32_t k = 0; k < count; ++k) out.bendCompliance[k] = gb[index[k]];
    index = mergeIndices(b, NUMI_CLOTH_HIP_MERGE_PARTICLE, specimen, &count);
    for (uint32_t k = 0; k < count; ++k) {
        if (out.inverseMass != nullptr) out.inverseMass[k] = gw[index[k]];
        if (out.damping != nullptr) out.damping[k] = gc[index[k]];
    }
}

}  // namespace

void numi_hip::batchSpecimenCounts(const NumiClothHipBatch& b, uint32_t specimen, uint32_t counts[3]) {
    if (specimen >= b.specimens) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "specimen out of range");
    mergeIndices(b, NUMI_CLOTH_HIP_MERGE_PARTICLE, specimen, &counts[0]);
    mergeIndices(b, NUMI_CLOTH_HIP_MERGE_DISTANCE, specimen, &counts[1]);
    mergeIndices(b, NUMI_CLOTH_HIP_MERGE_BEND, specimen, &counts[2]);
}

extern "C" {

int numi_cloth_hip_self_broad_phase(NumiClothHipSolver* solver, NumiClothHipSelfBroadPhase* io) {
    return guarded([&] {
        if (solver == nullptr || io == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        selfBroadPhase(*solver, io);
    });
}

int numi_cloth_hip_tile_info(NumiClothHipSolver* solver, NumiClothHipTileInfo* out) {
    return guarded([&] {
        if (solver == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        std::memset(out, 0, sizeof(*out));
        const auto& t = solver->tileEven;
        out->enabled = solver->useTiles && t.valid ? 1u : 0u;
        if (!out->enabled) return;
        out->tiles = t.tiles;
        out->ownedPerTile = t.owned;
        out->maxLoaded = t.maxLoaded;
        out->halo = t.halo;
        out->workgroupSize = t.block;
        out->passes = t.passes;
        out->staged = t.valid && t.staged ? 1u : 0u;
        out->redundancy = t.unique != 0 ? static_cast<double>(t.computed) / static_cast<double>(t.unique) : 0.0;
    });
}

int numi_cloth_hip_time_pass(NumiClothHipSolver* solver, uint32_t pass, uint32_t colour, uint32_t repeats,
                             NumiClothHipPassTiming* out) {
    return guarded([&] {
        if (solver == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        timePass(*solver, pass, colour, repeats, out);
    });
}


uint32_t numi_cloth_hip_abi_version(void) { return NUMI_CLOTH_BAG_GPU_ABI_VERSION; }

int numi_cloth_hip_create(const NumiClothHipSceneDesc* scene, NumiClothHipSolver** out) {
    return guarded([&] {
        if (scene == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto solver = std::make_unique<NumiClothHipSolver>();
        createSolver(*scene, *solver);
        *out = solver.release();
    });
}

int numi_cloth_hip_step(NumiClothHipSolver* solver, const NumiClothBagGPUConfig* configs, uint32_t substeps,
                        uint32_t iterations, uint32_t strainSweeps) {
    return guarded([&] {
        if (solver == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null solver");
        stepSolver(*solver, configs, substeps, iterations, strainSweeps);
    });
}

int numi_cloth_hip_step_noalloc(NumiClothHipSolver* solver, const NumiClothBagGPUConfig* configs,
                                uint32_t substeps, uint32_t iterations, uint32_t strainSweeps) {
    return guarded([&] {
        if (solver == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null solver");
        stepSolverNoAlloc(*solver, configs, substeps, iterations, strainSweeps);
    });
}

int numi_cloth_hip_read(NumiClothHipSolver* solver, const NumiClothHipState* out) {
    return guarded([&] {
        if (solver == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        readSolver(*solver, *out);
    });
}

int numi_cloth_hip_diagnostics(NumiClothHipSolver* solver, NumiClothHipDiagnostics* out) {
    return guarded([&] {
        if (solver == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        diagnoseSolver(*solver, out);
    });
}

int numi_cloth_hip_synchronize(NumiClothHipSolver* solver) {
    return guarded([&] {
        if (solver == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null solver");
        HIP_OK(hipStreamSynchronize(solver->stream));
    });
}

int numi_cloth_hip_launch_mode(NumiClothHipSolver* solver, NumiClothHipLaunchMode* out) {
    return guarded([&] {
        if (solver == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        out->graphs = solver->useGraphs ? 1u : 0u;
        out->decided = solver->launchDecided ? 1u : 0u;
        out->probeGraphMilliseconds = solver->launchProbeGraphMs;
        out->probeDirectMilliseconds = solver->launchProbeDirectMs;
    });
}

void* numi_cloth_hip_stream(NumiClothHipSolver* solver) {
    return solver == nullptr ? nullptr : static_cast<void*>(solver->stream);
}

int numi_cloth_hip_destroy(NumiClothHipSolver* solver) {
    return guarded([&] {
        if (solver == nullptr) return;
        (void)hipStreamSynchronize(solver->stream);
        delete solver;
    });
}

int numi_cloth_hip_batch_create(const NumiClothHipSceneDesc* const* specimens, uint32_t count,
                                NumiClothHipBatch** out) {
    return guarded([&] {
        if (specimens == nullptr || out == nullptr || count == 0) {
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument or no specimens");
        }
        *out = nullptr;
        auto batch = std::make_unique<NumiClothHipBatch>();
        if (numi_cloth_scene_merge(specimens, count, &batch->merged) != 0) {
            fail(NUMI_CLOTH_HIP_ERROR_TOPOLOGY, numi_cloth_hip_last_error());
        }
        batch->specimens = count;
        batch->solver = std::make_unique<NumiClothHipSolver>();
        createSolver(batch->merged->desc, *batch->solver);
        *out = batch.release();
    });
}

int numi_cloth_hip_batch_step(NumiClothHipBatch* batch, const NumiClothBagGPUConfig* configs, uint32_t substeps,
                              uint32_t iterations, uint32_t strainSweeps) {
    return guarded([&] {
        if (batch == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null batch");
        stepBatch(*batch, configs, substeps, iterations, strainSweeps);
    });
}

int numi_cloth_hip_batch_step_resident(NumiClothHipBatch* batch, const NumiClothBagGPUConfig* configs,
                                       uint32_t substeps, uint32_t iterations, uint32_t strainSweeps) {
    return guarded([&] {
        if (batch == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null batch");
        stepResident(*batch, configs, substeps, iterations, strainSweeps);
    });
}

int numi_cloth_hip_batch_read(NumiClothHipBatch* batch, uint32_t specimen, const NumiClothHipState* out) {
    return guarded([&] {
        if (batch == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        readBatch(*batch, specimen, *out);
    });
}

NumiClothHipSolver* numi_cloth_hip_batch_solver(NumiClothHipBatch* batch) {
    return batch == nullptr ? nullptr : batch->solver.get();
}

const NumiClothHipSceneDesc* numi_cloth_hip_batch_scene(const NumiClothHipBatch* batch) {
    return batch == nullptr ? nullptr : &batch->merged->desc;
}

int numi_cloth_hip_batch_adjoint(NumiClothHipBatch* batch, const NumiClothHipAdjointProblem* problem) {
    return guarded([&] {
        if (batch == nullptr || problem == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        runAdjoint(*batch, *problem);
    });
}

int numi_cloth_hip_batch_adjoint_stats(NumiClothHipBatch* batch, NumiClothHipAdjointStats* out) {
    return guarded([&] {
        if (batch == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        if (!batch->adjoint.ready) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "no adjoint has run on this batch");
        AdjointBuffers& A = batch->adjoint;
        uint32_t taped = 0;
        A.strainTaped.download(&taped, 1, batch->solver->stream);
        HIP_OK(hipStreamSynchronize(batch->solver->stream));
        out->strainPassesTaped = taped;
        out->strainPasses = static_cast<uint64_t>(A.lastSubsteps) * A.lastPasses * batch->specimens;
    });
}

int numi_cloth_hip_batch_gradients(NumiClothHipBatch* batch, uint32_t specimen,
                                   const NumiClothHipSpecimenGradients* out) {
    return guarded([&] {
        if (batch == nullptr || out == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        readGradients(*batch, specimen, *out);
    });
}

int numi_cloth_hip_batch_destroy(NumiClothHipBatch* batch) {
    return guarded([&] {
        if (batch != nullptr) {
            (void)hipStreamSynchronize(batch->solver->stream);
            delete batch;
        }
    });
}

}  // extern "C"
