/* Host shim so the device-side coder headers (mchip_device.hpp,
 * mchip_codec.hpp) compile as plain C++ for the CPU test harness.  Each
 * intrinsic reproduces the gfx950 instruction's result on the values the
 * coder feeds it. */
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define MCHIP_SCHED_FENCE() ((void)0)
struct uint2 { uint32_t x, y; };
struct uint4 { uint32_t x, y, z, w; };
static inline uint2 make_uint2(uint32_t x, uint32_t y) { uint2 r = {x, y}; return r; }
/* v_mul_u32_u24: low 32 bits of the product of the low 24 bits */
static inline uint32_t __umul24(uint32_t a, uint32_t b) {
    return (uint32_t)((uint64_t)(a & 0xFFFFFFu) * (uint64_t)(b & 0xFFFFFFu));
}
static inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline int __popc(uint32_t x) { return __builtin_popcount(x); }
/* a wave of one lane: bit 0 is the lane's predicate */
static inline uint64_t __builtin_amdgcn_ballot_w64(bool b) { return b ? 1u : 0u; }
static inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { uint4 r = {x, y, z, w}; return r; }
