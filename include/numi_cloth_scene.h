#pragma once

/*
 * Host-side scene construction and input artifacts of libnumi_cloth_hip.so.
 * No GPU is touched by anything declared here.
 *
 *  - Scene builders produce the ABI v13 arrays, greedy-coloured in
 *    construction order exactly as the reference colours its bag
 *    (reference tools/cloth_bag_metal.mm:627-676, :715-778, :833-892,
 *    :1047-1092): first free colour = lowest zero bit of the OR of the
 *    participants' 64-bit colour masks, then a stable sort by colour and
 *    one batch {first, count, colour, 0} per colour.
 *  - The material artifact parser follows reference
 *    include/numi/cloth_material.h:15-192 (schema numi.cloth.material.v1).
 *  - The grip trajectory loader/sampler follows reference
 *    include/numi/grip_trajectory.h:18-362 (schemas v1/v2/v3).
 */

#include <stddef.h>
#include <stdint.h>

#include "numi_cloth_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- material artifact (numi.cloth.material.v1) ---- */

/* The 19 parameters in the order of the reference's
 * ClothMaterialParameters (include/numi/cloth_material.h:18-38). */
typedef struct NumiClothMaterialValues {
    double ordinaryNodeMassKg;
    double hemNodeMassKg;
    double yarnRadiusM;
    double axialBodyComplianceMPerN;
    double axialCuffComplianceMPerN;
    double knotCompliance;
    double bendBodyComplianceMPerN;
    double bendCuffComplianceMPerN;
    double gripComplianceMPerN;
    double clothGroundFriction;
    double clothSelfFriction;
    double fruitYarnFriction;
    double fruitGroundFriction;
    double fruitPairFriction;
    double fruitRollingResistance;
    double yarnCrossflowDrag;
    double yarnSkinFriction;
    double fruitDrag;
    double fruitRotationalDrag;
} NumiClothMaterialValues;

typedef struct NumiClothMaterialArtifact {
    NumiClothMaterialValues values;
    uint32_t loaded;
    char parametersHash[128];
    char observationsHash[128];
} NumiClothMaterialArtifact;

int numi_cloth_material_defaults(NumiClothMaterialArtifact* out);
/* Strict parse; on failure returns NUMI_CLOTH_HIP_ERROR_PARSE with the
 * reference's message text in numi_cloth_hip_last_error(). */
int numi_cloth_material_load(const char* path, NumiClothMaterialArtifact* out);

/* ---- grip trajectory (numi.grip.trajectory.v1/v2/v3) ---- */

typedef struct NumiClothGripPose {
    double timeSeconds;
    double translation[3];
    double orientation[4]; /* xyzw */
    uint32_t active;
    uint32_t attachmentGeneration;
} NumiClothGripPose;

typedef struct NumiClothGripTrajectoryInfo {
    uint32_t poseCount;
    uint32_t selectNearestCuffPatch; /* 1 for schema v3 */
    uint32_t attachmentGenerations;
    uint32_t reserved;
    double durationSeconds;
    double maximumRotation;
    char schema[32];
    char contentFingerprint[32];     /* FNV-1a 64 of the file bytes, "0x%llx" */
} NumiClothGripTrajectoryInfo;

typedef struct NumiClothGripTrajectory NumiClothGripTrajectory;

int numi_cloth_grip_trajectory_load(const char* path,
                                    NumiClothGripTrajectory** out);
int numi_cloth_grip_trajectory_info(const NumiClothGripTrajectory* trajectory,
                                    NumiClothGripTrajectoryInfo* out);
int numi_cloth_grip_trajectory_pose(const NumiClothGripTrajectory* trajectory,
                                    uint32_t index, NumiClothGripPose* out);
int numi_cloth_grip_trajectory_sample(const NumiClothGripTrajectory* trajectory,
                                      double timeSeconds,
                                      NumiClothGripPose* out);
void numi_cloth_grip_trajectory_destroy(NumiClothGripTrajectory* trajectory);

/* ---- scene builders ---- */

typedef struct NumiClothHipScene NumiClothHipScene;

/* Square woven sheet in the z = elevation plane, knot (x, y) at index
 * y * width + x, centred on the origin. Distances are built row-major
 * horizontal then vertical; bends likewise; crossing knots optional. */
typedef struct NumiClothHipSheetSpec {
    uint32_t width;
    uint32_t height;
    float spacing;
    float elevation;
    float jitter;               /* z offsets uniform in [-jitter, jitter] */
    float arealDensity;         /* kg/m^2; knot mass = density * spacing^2 */
    float yarnRadius;
    float stretchCompliance;
    float bendCompliance;
    float knotCompliance;
    float strainLimit;
    float cellSize;             /* self-contact hash cell */
    float timestep;
    float gravityZ;
    float groundFriction;
    float selfFriction;
    uint32_t crossingKnots;     /* 1: add warp/weft crossing constraints */
    uint32_t ground;            /* 1: ground plane z = 0 active */
    uint32_t pinTopRow;         /* 1: row y = height-1 has zero inverse mass */
    uint32_t selfContact;       /* 1: hash self-contact */
    uint32_t sphereMode;        /* 0 none, 1 one fixed sphere, 2 random set */
    uint32_t sphereCount;       /* mode 2 */
    float sphereRadiusMin;      /* mode 2 */
    float sphereRadiusMax;      /* mode 2 */
    float fixedSphere[4];       /* mode 1: centre xyz, radius */
    uint64_t seed;
} NumiClothHipSheetSpec;

/* The reference's produce bag (reference tools/cloth_bag_metal.mm:33-38,
 * :420-440, :524-1274): 48 x 28 wall knots plus the 11 x 11 interior of a
 * 13 x 13 bottom grid, 12 fruits. Overrides <= 0 keep the material value. */
typedef struct NumiClothHipBagSpec {
    NumiClothMaterialValues material;
    float stretchComplianceOverride;
    float bendComplianceOverride;
    float strainLimitOverride;
    float fruitRadiusMin;       /* > 0: radii uniform from seed */
    float fruitRadiusMax;
    uint32_t selfContact;       /* 1: hash self-contact (replaces the
                                   reference's static all-pairs schedule) */
    uint32_t reserved;
    uint64_t seed;
} NumiClothHipBagSpec;

/* A hanging strip specimen (batched calibration and adjoint workloads): a
 * width x height knot grid in the x-z plane, top row (j = 0) pinned and
 * driven vertically by amplitude * (sin(2 pi frequency t + phase) -
 * sin(phase)) from where it starts, no ground, warp/weft distances and
 * chord bends. Knot (i, j) sits at
 * origin + (i spacing, jitter, -j spacing), row-major. */
typedef struct NumiClothHipStripSpec {
    uint32_t width;
    uint32_t height;
    float spacing;
    float arealDensity;         /* kg/m^2; knot mass = density * spacing^2 */
    float yarnRadius;
    float stretchCompliance;
    float bendCompliance;
    float strainLimit;
    float timestep;
    float gravityZ;
    float driveAmplitude;       /* m */
    float driveFrequency;       /* Hz */
    float drivePhase;           /* rad */
    float damping;              /* 1/s, see NumiClothHipSceneExtension */
    float origin[3];
    float jitter;               /* y offsets uniform in [-jitter, jitter] */
    uint64_t seed;
} NumiClothHipStripSpec;

int numi_cloth_scene_sheet(const NumiClothHipSheetSpec* spec,
                           NumiClothHipScene** out);
int numi_cloth_scene_strip(const NumiClothHipStripSpec* spec,
                           NumiClothHipScene** out);

/* Merges independent specimens into one scene that a single solver steps:
 * particles are concatenated in specimen order; the constraints of colour c
 * of every specimen form colour batch c of the merged scene (specimen order,
 * then each specimen's own batch order), so one colour launch covers all
 * specimens and every specimen evolves exactly as it would alone. Each
 * specimen's drive and damping become per-particle motion. All specimens
 * must share timestep, gravity, yarn radius and the ground flag, and have no
 * grips, fruits, mouth, static spheres, air loads or self-contact (those
 * couple particles beyond a specimen). */
int numi_cloth_scene_merge(const NumiClothHipSceneDesc* const* specimens,
                           uint32_t count, NumiClothHipScene** out);

enum NumiClothHipMergeFamily {
    NUMI_CLOTH_HIP_MERGE_PARTICLE = 0,
    NUMI_CLOTH_HIP_MERGE_DISTANCE = 1,
    NUMI_CLOTH_HIP_MERGE_BEND = 2,
    NUMI_CLOTH_HIP_MERGE_KNOT = 3
};

/* Where specimen `specimen`'s records of `family` went: *indices points at
 * *count merged indices, one per record of the specimen in its own order
 * (borrowed; valid until the merged scene is destroyed). */
int numi_cloth_scene_merge_indices(const NumiClothHipScene* merged, uint32_t family,
                                   uint32_t specimen, const uint32_t** indices,
                                   uint32_t* count);
int numi_cloth_scene_bag(const NumiClothHipBagSpec* spec,
                         NumiClothHipScene** out);
/* Borrowed view; valid until numi_cloth_scene_destroy. */
const NumiClothHipSceneDesc* numi_cloth_scene_desc(const NumiClothHipScene* scene);
void numi_cloth_scene_destroy(NumiClothHipScene* scene);

/* SplitMix64 (Steele, Lea, Flood 2014) used by every seeded builder. */
uint64_t numi_cloth_splitmix64(uint64_t* state);

/* FNV-1a 64 over raw bytes (reference tools/cloth_bag_metal.mm:4821-4860
 * hashes the raw FP32 buffers the same way). */
uint64_t numi_cloth_fnv1a64(const void* data, size_t bytes, uint64_t seed);

#ifdef __cplusplus
}
#endif
