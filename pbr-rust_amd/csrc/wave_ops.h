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
