#pragma once

/*
 * C ABI of libnumi_cloth_hip.so: the XPBD cloth substep loop on AMD Instinct
 * MI355X (gfx950).
 *
 * Every entry point takes plain pointers and sizes; no C++ or HIP type crosses
 * the boundary. Functions return 0 (NUMI_CLOTH_HIP_OK) or a positive status;
 * numi_cloth_hip_last_error() then holds a message for the calling thread.
 * Device-side failures are reported the way the reference reports them: as
 * the sticky NUMI_CLOTH_BAG_GPU_FAILURE_* word returned by read().
 *
 * The reference has no library; its de-facto host entry point is
 *   runGPU(device, queue, pipelines, InitialState, iterations, strainSweeps,
 *          perSubstepConfigs) -> GPUResult
 * (reference tools/cloth_bag_metal.mm:4010-4018, result :3951-3967).
 * create/step/read/destroy below split that call so that device state stays
 * resident between calls; step() over N substeps followed by step() over M
 * substeps gives the same bytes as one step() over N+M substeps.
 */

#include <stddef.h>
#include <stdint.h>

#include "numi/cloth_bag_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NUMI_CLOTH_HIP_EXTENSION_VERSION 2u

enum NumiClothHipStatus {
    NUMI_CLOTH_HIP_OK = 0,
    NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT = 1,
    NUMI_CLOTH_HIP_ERROR_TOPOLOGY = 2,
    NUMI_CLOTH_HIP_ERROR_DEVICE = 3,
    NUMI_CLOTH_HIP_ERROR_ALLOCATION = 4,
    NUMI_CLOTH_HIP_ERROR_UNSUPPORTED = 5,
    NUMI_CLOTH_HIP_ERROR_PARSE = 6,
    NUMI_CLOTH_HIP_ERROR_IO = 7      /* a file could not be opened or written */
};

/*
 * Scene facts that ABI v13 has no field for. Zero-initialise for the
 * reference behaviour (no static colliders, no kinematic drive, hash
 * self-contact off).
 */
typedef struct NumiClothHipSceneExtension {
    /* Static spheres: xyz centre, w radius. Cloth knots are pushed to
     * radius + yarn radius in the same pass as the ground. */
    const mr_float4* staticSpheres;
    uint32_t staticSphereCount;
    /* 1 enables yarn/yarn self-contact through the device uniform hash. */
    uint32_t selfContact;
    /* Broad-phase margin added to the swept segment bounds once per
     * substep; pairs beyond it are not tested again until the next substep. */
    float selfContactMargin;
    /* Maximum candidate pairs kept per substep (0 selects 8 per segment). */
    uint32_t selfCandidateCapacity;
    /* Kinematic drive of particles with zero inverse mass:
     * z(t) = anchor.z + amplitude * (sin(omega * t + phase) - sin(phase)),
     * anchor = the position given at create (where the knot starts),
     * t = time at the end of the substep. */
    float driveAmplitude;
    float driveOmega;
    float drivePhase;
    /* Linear velocity damping coefficient c >= 0 (1/s): at the start of a
     * substep v <- v / (1 + c dt) + g dt, so damping never scales the
     * gravity increment and static equilibria are unchanged.
     * 0 reproduces the reference. */
    float damping;
    /* Optional, one per particle: (drive amplitude, drive omega, drive
     * phase, damping). When set it replaces the four scalars above for
     * every particle, so independent specimens merged into one scene keep
     * their own drive and damping. NULL uses the scalars. */
    const mr_float4* particleMotion;
    /* Passes of the independent CPU reference's schedule (tools/cloth_bag.cpp),
     * absent from the Metal path; 0 (the default) leaves them out.
     * sweptGround: after the advance, with the ground on, a knot or fruit
     *   that started the substep at or above contact height and is
     *   predicted below it is put back at contact height; a fruit's removed
     *   advance / (w dt) is added to its ground impulse (reference
     *   sweepGroundPrediction, :2692-2721).
     * pointSelfCollision: in every iteration, after the segment self-contact
     *   pass, knots closer than two yarn radii that are not within two
     *   distance hops of each other are pushed apart by the reference's
     *   correction d * (2r - |d|) / (|d| (w_i + w_j)), shared by inverse mass
     *   (solveSelfCollision, :4249-4315). The reference applies the pairs one
     *   after another; here every pair acts on one position snapshot and each
     *   knot moves by the mean of its corrections (2^-32 m fixed-point sums,
     *   order independent), which is the reference's result for a knot in
     *   one contact. */
    uint32_t sweptGround;
    uint32_t pointSelfCollision;
    /* certificatePasses: at most this many extra certificate passes at the
     *   last substep of every step() call with strain sweeps, stopping after
     *   the first pass whose residuals all fall below the independent CPU
     *   reference's thresholds (its data-dependent loop at the last substep
     *   of a frame, up to 128 passes, tools/cloth_bag.cpp:5143-5185). A pass
     *   is one of the schedule's certificate rounds: segment self-contact,
     *   the strain sweeps, then the final contact passes (fruit pairs,
     *   fruit/yarn, ground). Residuals, each max(0, ...): strain-limit
     *   violation < 1e-6 m, ground penetration of knots and fruits < 1e-9 m,
     *   fruit/yarn penetration < 1e-6 m, segment self-penetration over the
     *   self-contact candidates < 1e-6 m. The passes run are reported by
     *   numi_cloth_hip_diagnostics. 0 (the default) runs none. */
    uint32_t certificatePasses;
} NumiClothHipSceneExtension;

/*
 * Everything the reference's InitialState carries
 * (reference tools/cloth_bag_metal.mm:466-488), as ABI v13 arrays.
 * Arrays of count 0 may be NULL.
 */
typedef struct NumiClothHipSceneDesc {
    const NumiClothBagGPUConfig* config;
    const NumiClothBagGPUParticle* particles;
    const NumiClothBagGPUDistance* distances;
    const NumiClothBagGPUBatch* distanceBatches;
    const NumiClothBagGPUGrip* grips;
    const NumiClothBagGPUKnot* knots;
    const NumiClothBagGPUBatch* knotBatches;
    const NumiClothBagGPUBend* bends;
    const NumiClothBagGPUBatch* bendBatches;
    const NumiClothBagGPUFruit* fruits;
    const NumiClothBagGPUFruitPair* fruitPairs;
    const NumiClothBagGPUBatch* fruitPairBatches;
    const NumiClothBagGPUYarnContact* yarnContacts;
    const uint32_t* mouthRimParticles;
    const uint32_t* cuffParticles;
    uint32_t particleCount;
    uint32_t distanceCount;
    uint32_t distanceBatchCount;
    uint32_t gripCount;
    uint32_t knotCount;
    uint32_t knotBatchCount;
    uint32_t bendCount;
    uint32_t bendBatchCount;
    uint32_t fruitCount;
    uint32_t fruitPairCount;
    uint32_t fruitPairBatchCount;
    uint32_t yarnContactCount;
    uint32_t mouthRimCount;
    uint32_t cuffCount;
    uint32_t initialCandidateMask;
    uint32_t initialReleasedMask;
    NumiClothHipSceneExtension extension;
} NumiClothHipSceneDesc;

/* Destination arrays for read(); any pointer may be NULL to skip it. */
typedef struct NumiClothHipState {
    NumiClothBagGPUParticle* particles;
    NumiClothBagGPUDistance* distances;
    NumiClothBagGPUGrip* grips;
    NumiClothBagGPUKnot* knots;
    NumiClothBagGPUBend* bends;
    NumiClothBagGPUFruit* fruits;
    NumiClothBagGPUFruitPair* fruitPairs;
    NumiClothBagGPUYarnContact* yarnContacts;
    NumiClothBagGPUSelfStatus* selfStatus;
    NumiClothBagGPUFrictionStatus* frictionStatus;
    NumiClothBagGPUAerodynamicsStatus* aerodynamicsStatus;
    NumiClothBagGPUReleaseStatus* releaseStatus;
    uint32_t* failure;
    /* Per fruit: largest mouth-release clearance seen since create, metres
     * (reference maximumMouthClearanceBits, tools/cloth_bag_metal.mm:7585-7591). */
    float* mouthClearance;
} NumiClothHipState;

/* Residual diagnostics, reduced on the device without float atomics
 * (replaces the host loops of reference tools/cloth_bag_metal.mm:5176-5276). */
typedef struct NumiClothHipDiagnostics {
    float maximumStrainViolation;   /* max(len - rest*(1+limit), 0) */
    float maximumGroundPenetration; /* max(r - z, 0) over knots */
    float maximumSpherePenetration; /* over static spheres */
    float maximumSelfPenetration;   /* over current self candidates */
    uint32_t selfCandidateCount;
    uint32_t selfCandidateOverflow;
    /* Fruit/yarn solves since create that had to rerun visit by visit (the
     * sparse schedule could not prove its skipped visits were no-ops), and
     * the most candidate visits one solve staged. */
    uint32_t yarnReruns;
    uint32_t yarnCandidatesMax;
    /* Self-contact broad phase of the last substep: (cell, segment) entries
     * the boxes requested and the grid's cell counts per axis. */
    uint32_t selfEntryCount;
    uint32_t selfGridCells[3];
    /* Knot pairs of the last point self-collision pass (0 when the scene
     * has none). */
    uint32_t pointSelfContacts;
    /* Certificate passes the last step() call ran (0 without
     * extension.certificatePasses). */
    uint32_t certificatePasses;
    /* Tile strain programs of the reconciliation rounds (and of the
     * certificate rounds of scenes without fruits or self-contact) since
     * create: tiles that returned at once because no tile owning one of
     * their knots changed them in the previous round (the round would
     * write the same bytes), and tiles that ran. */
    uint64_t tileRoundsSkipped;
    uint64_t tileRoundsRun;
    /* Present-mode fruit/yarn solves since create that tested every
     * (fruit, segment) visit and rebuilt the kept list of visits near
     * contact, and those that took their candidates from the kept list
     * (its knots and fruits had drifted too little for a visit off it to
     * reach contact). */
    uint32_t yarnListBuilds;
    uint32_t yarnListSolves;
} NumiClothHipDiagnostics;

typedef struct NumiClothHipSolver NumiClothHipSolver;

uint32_t numi_cloth_hip_abi_version(void);
const char* numi_cloth_hip_last_error(void);

/* Host->device boundary of runGPU (reference tools/cloth_bag_metal.mm:4019-4177):
 * validates topology (ranges, colour batches) on the host, converts the
 * records to device structure-of-arrays buffers and keeps them resident. */
int numi_cloth_hip_create(const NumiClothHipSceneDesc* scene,
                          NumiClothHipSolver** out);

/* The per-substep schedule of runGPU (reference
 * tools/cloth_bag_metal.mm:4211-4709): one substep per config. Configs may
 * change only what the reference lets a trajectory change (:4044-4090);
 * anything else returns NUMI_CLOTH_HIP_ERROR_TOPOLOGY. Asynchronous on the
 * solver's stream. */
int numi_cloth_hip_step(NumiClothHipSolver* solver,
                        const NumiClothBagGPUConfig* configs,
                        uint32_t substeps,
                        uint32_t iterations,
                        uint32_t strainSweeps);

/* step() for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in
 * the reference. Same results as numi_cloth_hip_step, with no hipMalloc,
 * hipFree or allocating library call: each call's configs and substep times
 * are staged in a fixed device array of the library (about 232 KB, held by
 * every device the library runs on until the process ends) and the substeps
 * are launched directly, not replayed from a captured graph. Limits: at most
 * 1024 substeps per call (more returns NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT),
 * and the strain-limit program for strainSweeps must exist (the solver
 * prepares 3 sweeps at create; one numi_cloth_hip_step call prepares any
 * other count), else NUMI_CLOTH_HIP_ERROR_UNSUPPORTED. Returns when the
 * configs are copied; the substeps run asynchronously on the solver's
 * stream. Calls on one device share the staging array: calls that overlap
 * on one device (several solvers stepped from several threads) race. */
int numi_cloth_hip_step_noalloc(NumiClothHipSolver* solver,
                                const NumiClothBagGPUConfig* configs,
                                uint32_t substeps,
                                uint32_t iterations,
                                uint32_t strainSweeps);

/* Device->host readback of runGPU (reference :4720-4757); synchronous. */
int numi_cloth_hip_read(NumiClothHipSolver* solver,
                        const NumiClothHipState* out);

int numi_cloth_hip_diagnostics(NumiClothHipSolver* solver,
                               NumiClothHipDiagnostics* out);

int numi_cloth_hip_synchronize(NumiClothHipSolver* solver);

/* Passes whose launches can be timed in isolation. */
enum NumiClothHipPass {
    NUMI_CLOTH_HIP_PASS_DISTANCE = 0,  /* one distance colour batch */
    NUMI_CLOTH_HIP_PASS_BEND = 1,      /* one bend colour batch */
    NUMI_CLOTH_HIP_PASS_STRAIN = 2,    /* one strain-limit colour batch */
    NUMI_CLOTH_HIP_PASS_COLLIDE = 3,   /* ground + static spheres over all knots */
    NUMI_CLOTH_HIP_PASS_TILE_ITERATION = 4 /* LDS tile program of an even iteration */
};

/* How the solver maps the substep onto the device (chosen at create). */
typedef struct NumiClothHipTileInfo {
    uint32_t enabled;          /* 1 when iterations run as LDS tile programs */
    uint32_t tiles;            /* workgroups per tile-program launch */
    uint32_t ownedPerTile;     /* knots written back per tile */
    uint32_t maxLoaded;        /* largest tile + halo held in LDS */
    uint32_t halo;             /* halo depth in constraint hops */
    uint32_t workgroupSize;
    uint32_t passes;           /* colour passes fused per even-iteration launch */
    uint32_t staged;           /* 1 when each tile's whole program is staged in LDS first */
    double redundancy;         /* constraint solves per launch / unique ones */
} NumiClothHipTileInfo;

int numi_cloth_hip_tile_info(NumiClothHipSolver* solver, NumiClothHipTileInfo* out);

typedef struct NumiClothHipPassTiming {
    double averageMilliseconds;        /* per launch, from HIP events */
    double algorithmicBytes;           /* unique bytes one launch must move */
    uint32_t launches;
    uint32_t itemsPerLaunch;
} NumiClothHipPassTiming;

/* Times `repeats` back-to-back launches of one pass on the solver's stream
 * (colour batch `colour` where the pass has colours) between two HIP events.
 * The launches act on the live state, as they would inside step(). */
int numi_cloth_hip_time_pass(NumiClothHipSolver* solver, uint32_t pass, uint32_t colour,
                             uint32_t repeats, NumiClothHipPassTiming* out);

/* Broad phase of the hash self-contact run on the live state (positions and
 * substep-start positions as they are now) with the scene's initial
 * configuration, for inspection. Outputs: the (cell id, segment) entries in
 * radix-sorted order, the candidate pairs (s < t) in emission order, and the
 * grid. Pointers may be NULL; counts are always written. */
typedef struct NumiClothHipSelfBroadPhase {
    uint32_t* pairs;           /* 2 x pairCapacity: s, t */
    uint32_t* cellIds;         /* entryCapacity */
    uint32_t* segments;        /* entryCapacity */
    uint32_t pairCapacity;
    uint32_t entryCapacity;
    uint32_t pairCount;        /* out: candidates found */
    uint32_t entryCount;       /* out: entries written */
    float origin[3];           /* out: grid origin */
    float cell;                /* out: cell size */
    uint32_t cells[3];         /* out: nx, ny, nz */
    uint32_t valid;            /* out: 0 if the grid was rejected */
} NumiClothHipSelfBroadPhase;

int numi_cloth_hip_self_broad_phase(NumiClothHipSolver* solver, NumiClothHipSelfBroadPhase* io);

/*
 * How the solver launches its substeps (no counterpart in the reference,
 * whose substeps are Metal command buffers): graphs = 1 replays a captured
 * hipGraph per substep, 0 launches the kernels one by one; the bytes are
 * the same. Unless NUMI_CLOTH_HIP_GRAPHS fixes it, the first step() of at
 * least five substeps times two substeps each way (after one replay that
 * uploads the graph) and keeps the faster; decided = 0 until then.
 */
typedef struct NumiClothHipLaunchMode {
    uint32_t graphs;
    uint32_t decided;
    float probeGraphMilliseconds;   /* two substeps by graph replay */
    float probeDirectMilliseconds;  /* two substeps by direct launches */
} NumiClothHipLaunchMode;

int numi_cloth_hip_launch_mode(NumiClothHipSolver* solver, NumiClothHipLaunchMode* out);

/* The hipStream_t all of the solver's work is queued on. */
void* numi_cloth_hip_stream(NumiClothHipSolver* solver);

int numi_cloth_hip_destroy(NumiClothHipSolver* solver);

/*
 * Batched specimens (no counterpart in the reference, which steps one scene
 * per runGPU call): many independent scenes merged into one device scene
 * (numi_cloth_scene_merge) and stepped together, so that every colour
 * launch, LDS tile program and hipGraph replay covers all specimens. Each
 * specimen evolves bit for bit as it would alone.
 */
typedef struct NumiClothHipBatch NumiClothHipBatch;

int numi_cloth_hip_batch_create(const NumiClothHipSceneDesc* const* specimens,
                                uint32_t count, NumiClothHipBatch** out);

/* configs[k] drives substep k of every specimen; its count fields are
 * replaced by the batch totals. */
int numi_cloth_hip_batch_step(NumiClothHipBatch* batch,
                              const NumiClothBagGPUConfig* configs,
                              uint32_t substeps, uint32_t iterations,
                              uint32_t strainSweeps);

/*
 * The same substeps as numi_cloth_hip_batch_step in one persistent launch
 * (no counterpart in the reference): one workgroup per specimen runs every
 * substep of the call with the specimen's knots and multipliers resident in
 * LDS, so HBM is touched only to load the state, read the constraint records
 * and store the state at the end. The batch ends bit for bit where
 * numi_cloth_hip_batch_step leaves it. For contact-free specimens
 * (distances, bends, strain limits, drive, damping) of at most 5120 knots
 * whose knots (16 B each) and multipliers (4 B per distance and bend) fit in
 * one workgroup's LDS; other batches fail with
 * NUMI_CLOTH_HIP_ERROR_UNSUPPORTED.
 */
int numi_cloth_hip_batch_step_resident(NumiClothHipBatch* batch,
                                       const NumiClothBagGPUConfig* configs,
                                       uint32_t substeps, uint32_t iterations,
                                       uint32_t strainSweeps);

/* Specimen `specimen`'s particles and constraint records in its own order and
 * indexing; failure and friction status are those of the whole batch. */
int numi_cloth_hip_batch_read(NumiClothHipBatch* batch, uint32_t specimen,
                              const NumiClothHipState* out);

/* Borrowed solver of the merged scene (synchronize, diagnostics, tile_info,
 * time_pass, stream) and its scene; valid until batch_destroy. */
NumiClothHipSolver* numi_cloth_hip_batch_solver(NumiClothHipBatch* batch);
const NumiClothHipSceneDesc* numi_cloth_hip_batch_scene(const NumiClothHipBatch* batch);

/*
 * Adjoint of a batch trajectory (no counterpart in the reference, whose
 * calibration fitter takes finite differences of whole runs,
 * tools/cloth_calibration.cpp:568-650). For contact-free specimens
 * (distances, bends, strain limits, damping, driven pinned knots) it runs
 * `substeps` substeps forward from the batch's current state (the state
 * advances exactly as numi_cloth_hip_batch_step would leave it, through the
 * specimen kernels: one workgroup per specimen, knots and multipliers in
 * LDS), storing (x, v) every `checkpointInterval` substeps; then, segment by
 * segment from the last, recomputes the segment from its checkpoint with a
 * tape of every pass and runs it backwards. The loss of specimen s is
 *   NUMI_CLOTH_ADJOINT_TRACKING:
 *     L_s = sum over substeps t and tracked knots k of |x_k(t) - target_k(t)|^2
 *   NUMI_CLOTH_ADJOINT_LINEAR (an observable for the calibration fitter,
 *   e.g. a knot's height after the last substep):
 *     L_s = sum over substeps t and tracked knots k of target_k(t) . x_k(t)
 *     (`targets` holds the weights)
 * with x_k(t) the position after substep t.
 */
#define NUMI_CLOTH_ADJOINT_TRACKING 0u
#define NUMI_CLOTH_ADJOINT_LINEAR 1u
typedef struct NumiClothHipAdjointProblem {
    const NumiClothBagGPUConfig* configs;  /* one per substep; counts replaced */
    uint32_t substeps;
    uint32_t iterations;
    uint32_t strainSweeps;
    uint32_t checkpointInterval;           /* substeps per checkpoint segment */
    const uint32_t* tracked;               /* specimens x trackedPerSpecimen, specimen-local knots */
    uint32_t trackedPerSpecimen;
    uint32_t observable;                   /* NUMI_CLOTH_ADJOINT_TRACKING or _LINEAR */
    const float* targets;                  /* substeps x specimens x trackedPerSpecimen x 3 */
} NumiClothHipAdjointProblem;

/* Gradients of one specimen's loss, in the specimen's own record order:
 * dL/d(compliance) per distance and per bend, dL/dw per knot (inverse mass;
 * for an areal density rho, dL/drho = -sum_i dL/dw_i * w_i / rho over the
 * free knots) and dL/dc per knot for the damping coefficient c. */
typedef struct NumiClothHipSpecimenGradients {
    double* loss;
    float* distanceCompliance;
    float* bendCompliance;
    float* inverseMass;
    float* damping;
} NumiClothHipSpecimenGradients;

int numi_cloth_hip_batch_adjoint(NumiClothHipBatch* batch, const NumiClothHipAdjointProblem* problem);
/* What the last adjoint call taped (no counterpart in the reference): of
 * the strain-limit passes its recomputations ran (substeps x strain passes
 * per substep x specimens), those in which a constraint acted and whose
 * positions were therefore taped and reversed. */
typedef struct NumiClothHipAdjointStats {
    uint64_t strainPasses;
    uint32_t strainPassesTaped;
    uint32_t reserved;
} NumiClothHipAdjointStats;

int numi_cloth_hip_batch_adjoint_stats(NumiClothHipBatch* batch, NumiClothHipAdjointStats* out);

int numi_cloth_hip_batch_gradients(NumiClothHipBatch* batch, uint32_t specimen,
                                   const NumiClothHipSpecimenGradients* out);

int numi_cloth_hip_batch_destroy(NumiClothHipBatch* batch);

#ifdef __cplusplus
}
#endif
