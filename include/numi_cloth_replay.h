#ifndef NUMI_CLOTH_REPLAY_H
#define NUMI_CLOTH_REPLAY_H

/* Frame driver of the produce-bag trajectory replays: grip pose generators,
 * the per-frame substep configs, the scenario's initial state, FNV-1a frame
 * hashes, the largest handle lag and OBJ dumps for the renderer.
 *
 * Replaces the reference's host frame loop runTrajectoryReplay
 * (tools/cloth_bag_metal.mm:5050-5174) with its helpers trajectoryGripPose
 * (:5079-5125), pickupGripTarget / spinGripTarget (:361-388),
 * makeTrajectoryInitialState (:5127-5164), hashGPUResult (:4822-4860) and
 * dumpGPUOBJ (:4862-4938). The reference calls runGPU once per frame and
 * re-packs the result as the next frame's input (continuedInitialState,
 * :4764-4782); here the solver keeps the state in device memory between
 * frames. */

#include <stdint.h>

#include "numi_cloth_hip.h"
#include "numi_cloth_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum NumiClothReplayScenario {
    NUMI_CLOTH_REPLAY_GROUNDED = 0, /* grip inactive at the cuff, ground on */
    NUMI_CLOTH_REPLAY_SPIN = 1,     /* bag lifted 0.52 m, ground off, 0.28 m circle at 4.8 rad/s */
    NUMI_CLOTH_REPLAY_PICKUP = 2,   /* lift 1.25 m, snap at 1.0 s, recovery at 1.45 s */
    NUMI_CLOTH_REPLAY_RECORDED = 3  /* a numi.grip.trajectory file */
} NumiClothReplayScenario;

/* Substeps per frame and substep length of the replays (reference
 * kPickupSubstepsPerFrame, kTimestep: tools/cloth_bag_metal.mm:80, :134). */
#define NUMI_CLOTH_REPLAY_SUBSTEPS_PER_FRAME 48u
#define NUMI_CLOTH_REPLAY_TIMESTEP (1.0f / 5760.0f)

typedef struct NumiClothReplayPose {
    double target[3];        /* world grip centre, m */
    double orientation[4];   /* xyzw */
    uint32_t active;
    uint32_t attachmentGeneration;
} NumiClothReplayPose;

/* Grip pose of a scenario at time t (s). The base point is the authored
 * position of cuff knot (level 27, ring 0), lifted 0.52 m for spin; the
 * recorded scenario adds the trajectory's sampled translation. trajectory
 * may be NULL except for NUMI_CLOTH_REPLAY_RECORDED. */
int numi_cloth_replay_pose(NumiClothReplayScenario scenario, double timeSeconds,
                           const NumiClothGripTrajectory* trajectory, NumiClothReplayPose* out);

/* Turns a bag scene (numi_cloth_scene_bag) into the scenario's initial
 * state in place: spin lifts every knot and fruit (position and previous
 * position) by 0.52 m and turns the ground off, the others turn it on; the
 * config's grip fields take the pose at t = 0 (target, active,
 * orientation, attachment generation, nearest-cuff-patch flag of a v3
 * trajectory). */
int numi_cloth_replay_prepare(NumiClothHipScene* scene, NumiClothReplayScenario scenario,
                              const NumiClothGripTrajectory* trajectory);

/* The substep configs of frame `frame` (0-based): each is `base` with the
 * grip fields of the pose at t = (frame * substepsPerFrame + s + 1) * dt,
 * dt = NUMI_CLOTH_REPLAY_TIMESTEP, s = 0 .. substepsPerFrame - 1. */
int numi_cloth_replay_frame_configs(const NumiClothBagGPUConfig* base, NumiClothReplayScenario scenario,
                                    const NumiClothGripTrajectory* trajectory, uint32_t frame,
                                    uint32_t substepsPerFrame, NumiClothBagGPUConfig* out);

/* FNV-1a 64 of a state in the reference's order (hashGPUResult): the raw
 * bytes of particles, distances, grips, knots, bends, fruits, fruit pairs,
 * yarn contacts, then the self, friction, aerodynamics and release status
 * records, the per-fruit mouth clearance bits, and finally the failure
 * word xored in whole. Record counts come from `scene`; the state's
 * pointers must be set for every non-empty family and for the status
 * records (NumiClothHipState as filled by numi_cloth_hip_read). */
int numi_cloth_state_hash(const NumiClothHipSceneDesc* scene, const NumiClothHipState* state, uint64_t* out);

/* Wavefront OBJ of a bag state for the renderer (reference dumpGPUOBJ):
 * "# vertices" / render-triangle header, one "v" line per knot, two
 * triangles per wall quad and per bottom-grid quad (diagonal alternating
 * with row + column parity), a "# grip center ..." line from `config` and
 * `grips`, and one "# ball ..." line per fruit. Numbers are written with 9
 * significant digits. Bag topology only (1,465 knots). */
int numi_cloth_obj_write(const char* path, const NumiClothBagGPUParticle* particles, uint32_t particleCount,
                         const NumiClothBagGPUFruit* fruits, uint32_t fruitCount, const NumiClothBagGPUGrip* grips,
                         uint32_t gripCount, const NumiClothBagGPUConfig* config);

typedef struct NumiClothReplayResult {
    uint64_t* frameHashes;       /* caller array of `frames` entries (may be NULL) */
    uint32_t framesRun;          /* frames stepped (stops after the first failing frame) */
    uint32_t failureFree;        /* 1 when no frame raised a failure bit */
    uint32_t releasedMask;       /* released-fruit mask after the last frame */
    uint32_t reserved;
    double maximumHandleLag;     /* largest |grip knot - grip target| at a frame's end, m */
    double stepSeconds;          /* wall time of the frames' step calls, each synchronized */
} NumiClothReplayResult;

/* Steps `frames` frames of a scenario on a solver created from the
 * prepared scene (base = that scene's config), reading the state back after
 * every frame to hash it and to measure the handle lag (grip scenarios
 * while the grip is active), as the reference's frame loop does. When
 * dumpPrefix is not NULL, writes "<prefix>-0.obj" from `scene` before the
 * first frame and "<prefix>-<n>.obj" after frame n when n is a multiple of
 * dumpEvery or the last frame. */
int numi_cloth_replay_run(NumiClothHipSolver* solver, const NumiClothHipSceneDesc* scene,
                          NumiClothReplayScenario scenario, const NumiClothGripTrajectory* trajectory,
                          uint32_t frames, uint32_t iterations, uint32_t strainSweeps, uint32_t dumpEvery,
                          const char* dumpPrefix, NumiClothReplayResult* out);

#ifdef __cplusplus
}
#endif

#endif
