// Wave64 / workgroup helpers for gfx950 (CDNA4): every wavefront is 64 lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbrtm {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint32_t wave_inclusive_add(uint32_t v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t o = __shfl_up(v, d, kWave);
        if (__lane_id() >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = min(v, (uint32_t)__shfl_xor(v, d, kWave));
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, kWave));
    return v;
}

// Reductions for code where all 64 lanes are active and every lane needs the
// result: DPP exchanges within rows of 16 lanes (quad permutes, half-row and row
// mirrors), then the four row results read as scalars. No LDS traffic.
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce_rows(uint32_t v, Op op) {
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0xB1, 0xF, 0xF, false)));   // lanes 1,0,3,2
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x4E, 0xF, 0xF, false)));   // lanes 2,3,0,1
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x141, 0xF, 0xF, false)));  // half-row mirror
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x140, 0xF, 0xF, false)));  // row mirror
    const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}
__device__ __forceinline__ uint32_t wave_min_u32_all(uint32_t v) {
    return wave_reduce_rows(v, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
}
__device__ __forceinline__ uint32_t wave_max_u32_all(uint32_t v) {
    return wave_reduce_rows(v, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
}
__device__ __forceinline__ uint32_t wave_sum_all(uint32_t v) {
    return wave_reduce_rows(v, [](uint32_t a, uint32_t b) { return a + b; });
}

// Exclusive scan across a workgroup of BLOCK threads; returns the thread's
// exclusive prefix and writes the block total to *total (all threads).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_add(uint32_t v, uint32_t *s_waves, uint32_t *total) {
    constexpr int kWaves = BLOCK / kWave;
    const int w = threadIdx.x / kWave;
    uint32_t inc = wave_inclusive_add(v);
    if (__lane_id() == kWave - 1) s_waves[w] = inc;
    __syncthreads();
    uint32_t wave_base = 0, sum = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        uint32_t t = s_waves[i];
        wave_base += (i < w) ? t : 0u;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return wave_base + inc - v;
}

}  // namespace pbrtm
