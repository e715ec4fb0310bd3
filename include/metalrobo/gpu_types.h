#pragma once

/*
 * 16-byte-aligned 4 x 32-bit records shared by host C/C++ and HIP device code.
 *
 * Interface parity: these are the vector record names the cloth ABI is written
 * in (reference include/metalrobo/gpu_types.h:1-32). Only the cloth subset is
 * declared here; the robot-model records of that header are not part of the
 * cloth path. Layout: four consecutive 32-bit lanes x,y,z,w, 16-byte aligned,
 * which is also the layout of HIP's float4 / uint4, so device code may
 * reinterpret these records as native vector loads.
 */

#include <stdint.h>

#if defined(__cplusplus)
#define MR_ALIGN16 alignas(16)
#else
#define MR_ALIGN16 _Alignas(16)
#endif

typedef uint32_t mr_u32;
typedef int32_t mr_i32;
typedef uint64_t mr_u64;

typedef struct MR_ALIGN16 mr_float4 {
    float x;
    float y;
    float z;
    float w;
} mr_float4;

typedef struct MR_ALIGN16 mr_uint4 {
    uint32_t x;
    uint32_t y;
    uint32_t z;
    uint32_t w;
} mr_uint4;
