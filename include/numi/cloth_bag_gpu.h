#pragma once

/*
 * Pointer-free cloth records, ABI version 13.
 *
 * Interface parity: every record name, field order and byte size matches the
 * reference ABI (include/numi/cloth_bag_gpu.h:5-209) so that arrays produced
 * by the reference harness can be handed to this library unchanged. The sizes
 * are asserted at the bottom of the file. On the device the solver does not
 * keep these records: it converts them once into structure-of-arrays buffers
 * and converts back only when the caller reads state.
 */

#include "metalrobo/gpu_types.h"

#define NUMI_CLOTH_BAG_GPU_ABI_VERSION 13u
#define NUMI_CLOTH_BAG_GPU_INVALID_PARTICLE 0xffffffffu
#define NUMI_CLOTH_BAG_GPU_SELF_IMPULSE_CAPACITY 4096u
#define NUMI_CLOTH_BAG_GPU_MOUTH_RIM_CAPACITY 64u

/* Sticky failure bits, OR-ed into one device word. */
enum NumiClothBagGPUFailure {
    NUMI_CLOTH_BAG_GPU_FAILURE_NONE = 0u,
    NUMI_CLOTH_BAG_GPU_FAILURE_ABI = 1u,
    NUMI_CLOTH_BAG_GPU_FAILURE_RANGE = 2u,
    NUMI_CLOTH_BAG_GPU_FAILURE_BATCH = 4u,
    NUMI_CLOTH_BAG_GPU_FAILURE_NONFINITE = 8u,
    NUMI_CLOTH_BAG_GPU_FAILURE_GRIP_CAPTURE = 16u
};

/* Per-substep configuration (224 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUConfig {
    mr_uint4 control;                 /* ABI, particles, distances, grips */
    mr_uint4 constraintCounts;        /* knots, bends, ground on/off, fruits */
    mr_uint4 contactCounts;           /* fruit pairs, fruit/yarn records,
                                         static self pairs, self batches */
    mr_float4 gravityAndTimestep;     /* gravity xyz, substep dt */
    mr_float4 gripTargetAndActive;    /* handle position, active flag */
    mr_float4 gripOrientation;        /* handle quaternion xyzw */
    mr_uint4 gripControl;             /* generation, selection mode,
                                         knots per cuff row, patch half width */
    mr_float4 gripMaterial;           /* capture radius, reserved x3 */
    mr_float4 clothMaterial;          /* yarn radius, self cell size,
                                         ground friction, self friction */
    mr_float4 fruitMaterial;          /* pair, ground, rolling, yarn friction */
    mr_float4 airVelocityAndDensity;  /* air velocity xyz, density */
    mr_float4 aerodynamicCoefficients;/* yarn crossflow, yarn skin,
                                         fruit drag, fruit rotational drag */
    mr_uint4 mouthControl;            /* rim count, bottom centre, reserved */
    mr_float4 mouthMaterial;          /* release hysteresis, reserved */
} NumiClothBagGPUConfig;

/* Yarn knot (48 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUParticle {
    mr_float4 positionAndInverseMass; /* x, inverse mass */
    mr_float4 previousAndMass;        /* substep start position, mass */
    mr_float4 velocity;               /* v, predicted vz / ground impulse */
} NumiClothBagGPUParticle;

/* Yarn segment / distance constraint (32 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUDistance {
    mr_uint4 particlesAndColor;       /* i, j, colour, yarn kind */
    mr_float4 material;               /* rest, compliance, lambda,
                                         strain limit (relative) */
} NumiClothBagGPUDistance;

/* Handle attachment (48 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUGrip {
    mr_uint4 particle;                /* particle, generation,
                                         max capture distance bits, centre ring */
    mr_float4 targetOffsetAndCompliance;
    mr_float4 lambda;                 /* vector lambda, max capture error */
} NumiClothBagGPUGrip;

/* Warp/weft crossing-angle constraint (48 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUKnot {
    mr_uint4 particles;               /* warp a, warp b, weft a, weft b */
    mr_uint4 control;                 /* colour, reserved */
    mr_float4 material;               /* rest cosine, compliance, lambda */
} NumiClothBagGPUKnot;

/* Chord bending constraint (32 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUBend {
    mr_uint4 particlesAndColor;       /* first, middle, third, colour */
    mr_float4 material;               /* rest chord, rest arc,
                                         compliance, lambda */
} NumiClothBagGPUBend;

/* Rigid sphere ("fruit", 96 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUFruit {
    mr_float4 positionAndInverseMass;
    mr_float4 previousAndRadius;
    mr_float4 velocityAndGroundImpulse;
    mr_float4 angularVelocity;
    mr_float4 orientation;
    mr_uint4 identity;
} NumiClothBagGPUFruit;

/* Sphere/sphere contact (32 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUFruitPair {
    mr_uint4 fruitsAndColor;          /* a, b, colour, stable index */
    mr_float4 contact;                /* impulse-weighted normal, impulse */
} NumiClothBagGPUFruitPair;

/* Sphere/yarn-segment contact record (112 bytes). */
typedef struct MR_ALIGN16 NumiClothBagGPUYarnContact {
    mr_uint4 identity;                /* fruit, segment, knot a, knot b */
    mr_float4 currentNormalAndSeparation;
    mr_float4 sweptNormalAndAdvance;
    mr_float4 weightsAndTime;         /* current w, swept w, impact t,
                                         combined radius */
    mr_uint4 control;                 /* overlap, impact, degenerate,
                                         accepted responses */
    mr_float4 fruitNormalAndImpulse;
    mr_float4 segmentImpulse;         /* knot a/b weights, first impact t,
                                         swept advance */
} NumiClothBagGPUYarnContact;

/* Static yarn/yarn pair of the reference's all-pairs schedule (8 bytes). */
typedef struct NumiClothBagGPUSelfPair {
    mr_u32 firstSegment;
    mr_u32 secondSegment;
} NumiClothBagGPUSelfPair;

typedef struct MR_ALIGN16 NumiClothBagGPUSelfStatus {
    mr_uint4 counters;                /* present, swept, max correction bits */
} NumiClothBagGPUSelfStatus;

typedef struct MR_ALIGN16 NumiClothBagGPUSelfImpulse {
    mr_uint4 identity;                /* pair, segment a, segment b, epoch */
    mr_float4 normalAndImpulse;
    mr_float4 endpointImpulses;
} NumiClothBagGPUSelfImpulse;

typedef struct MR_ALIGN16 NumiClothBagGPUFrictionStatus {
    mr_uint4 counters;                /* pair, yarn, cloth/ground, fruit/ground */
    mr_uint4 metrics;                 /* cone ratio bits, rolling ratio bits,
                                         rolling count, yarn/yarn count */
} NumiClothBagGPUFrictionStatus;

typedef struct MR_ALIGN16 NumiClothBagGPUAerodynamicsStatus {
    mr_uint4 metrics;                 /* yarn force, fruit force,
                                         fruit torque bits */
} NumiClothBagGPUAerodynamicsStatus;

typedef struct MR_ALIGN16 NumiClothBagGPUReleaseStatus {
    mr_uint4 masks;                   /* candidate mask, released mask */
} NumiClothBagGPUReleaseStatus;

typedef struct MR_ALIGN16 NumiClothBagGPUBatch {
    mr_uint4 control;                 /* first, count, colour, reserved */
} NumiClothBagGPUBatch;

#if defined(__cplusplus)
static_assert(sizeof(NumiClothBagGPUConfig) == 224, "config ABI");
static_assert(sizeof(NumiClothBagGPUParticle) == 48, "particle ABI");
static_assert(sizeof(NumiClothBagGPUDistance) == 32, "distance ABI");
static_assert(sizeof(NumiClothBagGPUGrip) == 48, "grip ABI");
static_assert(sizeof(NumiClothBagGPUKnot) == 48, "knot ABI");
static_assert(sizeof(NumiClothBagGPUBend) == 32, "bend ABI");
static_assert(sizeof(NumiClothBagGPUFruit) == 96, "fruit ABI");
static_assert(sizeof(NumiClothBagGPUFruitPair) == 32, "fruit pair ABI");
static_assert(sizeof(NumiClothBagGPUYarnContact) == 112, "yarn contact ABI");
static_assert(sizeof(NumiClothBagGPUSelfPair) == 8, "self pair ABI");
static_assert(sizeof(NumiClothBagGPUSelfStatus) == 16, "self status ABI");
static_assert(sizeof(NumiClothBagGPUSelfImpulse) == 48, "self impulse ABI");
static_assert(sizeof(NumiClothBagGPUFrictionStatus) == 32, "friction ABI");
static_assert(sizeof(NumiClothBagGPUAerodynamicsStatus) == 16, "aero ABI");
static_assert(sizeof(NumiClothBagGPUReleaseStatus) == 16, "release ABI");
static_assert(sizeof(NumiClothBagGPUBatch) == 16, "batch ABI");
#endif
