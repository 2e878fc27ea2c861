// Onesweep LSD radix sort (Adinets & Merrill 2022) for 30-bit Morton keys with
// u32 payloads, written for wave64: three 10-bit digit passes, each a single
// kernel that ranks a tile of 4096 keys (12288 from 4M keys up) in LDS
// (wave-level match by ballots), looks
// back across tiles for the per-digit global prefix (decoupled look-back over
// one 32-bit status word per tile and digit), and scatters through LDS.
// Stable: equal keys keep their input order, so ties stay in index order.
#include "pbrtm_internal.h"
#include "wave_ops.h"

namespace pbrtm {

namespace {

constexpr int kSortBlock = 512;
constexpr int kSortWaves = kSortBlock / kWave;
// Keys per thread: 8 (4096-key tiles) below kSortLargeN keys, 24 (12288-key tiles,
// 136 KB of LDS, one workgroup per CU) from there up, where three times fewer
// tiles shorten the look-back chains and the per-tile fixed work (10M keys:
// 0.66 -> 0.57 ms for the three passes).
constexpr int kSortItemsSmall = 8;
constexpr int kSortItemsLarge = 24;
constexpr uint32_t kSortLargeN = 1u << 22;
constexpr uint32_t sort_items(uint32_t n) { return n >= kSortLargeN ? kSortItemsLarge : kSortItemsSmall; }
constexpr uint32_t sort_tiles(uint32_t n) {
    return uint32_t((uint64_t(n) + uint64_t(kSortBlock) * sort_items(n) - 1) / (uint64_t(kSortBlock) * sort_items(n)));
}
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 24;

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Exclusive offsets of each digit for all passes: offsets[p][d] = sum hist[p][<d].
__global__ __launch_bounds__(kSortRadix) void digit_offsets_kernel(const uint32_t *__restrict__ hist,
                                                                   uint32_t *__restrict__ offsets) {
    __shared__ uint32_t s_waves[kSortRadix / kWave];
    for (uint32_t p = 0; p < kSortPasses; ++p) {
        uint32_t v = hist[p * kSortRadix + threadIdx.x];
        uint32_t tot;
        uint32_t ex = block_exclusive_add<kSortRadix>(v, s_waves, &tot);
        offsets[p * kSortRadix + threadIdx.x] = ex;
    }
}

__device__ __forceinline__ uint32_t status_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int kSortItems>
__global__ __launch_bounds__(kSortBlock) void onesweep_pass_kernel(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, uint32_t n, uint32_t shift,
    const uint32_t *__restrict__ digit_offsets, uint32_t *status, uint32_t *tile_counter,
    uint32_t *err) {
    constexpr int kSortTile = kSortBlock * kSortItems;
    __shared__ uint32_t s_hist[kSortWaves][kSortRadix];  // per-wave digit counts, then offsets
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_start[kSortRadix];  // digit start inside the sorted tile
    __shared__ uint32_t s_gbase[kSortRadix];  // digit start in the output
    __shared__ uint32_t s_waves[kSortWaves];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const int l = __lane_id();
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);  // launch-order tile id: look-back only waits on started tiles
    for (int i = tid; i < kSortWaves * kSortRadix; i += kSortBlock) (&s_hist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = uint64_t(tile) * kSortTile;

    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + uint64_t(w) * (kSortItems * kWave) + uint64_t(k) * kWave + l;
        const bool valid = idx < n;
        key[k] = valid ? keys_in[idx] : 0u;
        val[k] = valid ? vals_in[idx] : 0u;
        const uint32_t d = (key[k] >> shift) & (kSortRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < kSortRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : 0;
        const uint32_t cnt = __popcll(peers);
        const uint32_t below = __popcll(peers & lanemask_lt());
        uint32_t prior = 0;
        if (valid && l == leader) {
            prior = s_hist[w][d];
            s_hist[w][d] = prior + cnt;
        }
        prior = __shfl(prior, leader, kWave);
        rank[k] = valid ? prior + below : 0u;
    }
    __syncthreads();

    // Per digit: exclusive offsets across waves; tile count. Thread owns digits 2t, 2t+1.
    uint32_t cnt2[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int d = 2 * tid + j;
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ++ww) {
            uint32_t c = s_hist[ww][d];
            s_hist[ww][d] = run;
            run += c;
        }
        cnt2[j] = run;
        status_store(&status[uint64_t(tile) * kSortRadix + d], (tile == 0 ? kFlagPrefix : kFlagAgg) | run);
    }
    uint32_t tot;
    const uint32_t ex = block_exclusive_add<kSortBlock>(cnt2[0] + cnt2[1], s_waves, &tot);
    s_start[2 * tid] = ex;
    s_start[2 * tid + 1] = ex + cnt2[0];

#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int d = 2 * tid + j;
        uint32_t prefix = 0;
        if (tile > 0) {
            int64_t t = int64_t(tile) - 1;
            uint32_t spins = 0;
            while (t >= 0) {
                const uint32_t v = status_load(&status[uint64_t(t) * kSortRadix + d]);
                const uint32_t flag = v & ~kCountMask;
                if (flag == 0u) {
                    if (++spins > kSpinLimit) {
                        atomicOr(err, uint32_t(kErrSpin));
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += v & kCountMask;
                if (flag == kFlagPrefix) break;
                --t;
            }
            status_store(&status[uint64_t(tile) * kSortRadix + d], kFlagPrefix | (prefix + cnt2[j]));
        }
        s_gbase[d] = digit_offsets[d] + prefix;
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint64_t idx = base + uint64_t(w) * (kSortItems * kWave) + uint64_t(k) * kWave + l;
        if (idx < n) {
            const uint32_t d = (key[k] >> shift) & (kSortRadix - 1);
            const uint32_t pos = s_start[d] + s_hist[w][d] + rank[k];
            s_keys[pos] = key[k];
            s_vals[pos] = val[k];
        }
    }
    __syncthreads();
    const uint32_t tile_n = uint32_t(min(uint64_t(kSortTile), uint64_t(n) - base));
    for (uint32_t i = tid; i < tile_n; i += kSortBlock) {
        const uint32_t k = s_keys[i];
        const uint32_t d = (k >> shift) & (kSortRadix - 1);
        const uint32_t out = s_gbase[d] + (i - s_start[d]);
        if (out < n) {
            keys_out[out] = k;
            vals_out[out] = s_vals[i];
        } else {
            atomicOr(err, uint32_t(kErrSpin));
        }
    }
}

}  // namespace

hipError_t onesweep_reserve(SortScratch &s, uint32_t n) {
    const uint32_t tiles = sort_tiles(n);
    const uint64_t status_words = uint64_t(tiles) * kSortRadix * kSortPasses;
    hipError_t e = s.status.reserve(status_words * sizeof(uint32_t));
    if (e == hipSuccess) e = s.offsets.reserve(kSortPasses * kSortRadix * sizeof(uint32_t));
    if (e == hipSuccess) e = s.counters.reserve(64 * sizeof(uint32_t));
    return e;
}

hipError_t onesweep_sort_30(uint32_t *keys0, uint32_t *vals0, uint32_t *keys1, uint32_t *vals1,
                            uint32_t n, const uint32_t *hist, uint32_t *err, SortScratch &s,
                            hipStream_t stream, uint32_t **keys_result, uint32_t **vals_result) {
    return onesweep_sort(keys0, vals0, keys1, vals1, n, hist, err, s, stream, keys_result, vals_result, kSortPasses);
}

hipError_t onesweep_sort(uint32_t *keys0, uint32_t *vals0, uint32_t *keys1, uint32_t *vals1, uint32_t n,
                         const uint32_t *hist, uint32_t *err, SortScratch &s, hipStream_t stream,
                         uint32_t **keys_result, uint32_t **vals_result, uint32_t passes) {
    if (passes == 0 || passes > kSortPasses) return hipErrorInvalidValue;
    *keys_result = keys0;
    *vals_result = vals0;
    if (n == 0) return hipSuccess;
    const uint32_t tiles = sort_tiles(n);
    const uint64_t status_words = uint64_t(tiles) * kSortRadix * kSortPasses;
    hipError_t e = onesweep_reserve(s, n);
    if (e != hipSuccess) return e;
    uint32_t *status = s.status.as<uint32_t>();
    uint32_t *offsets = s.offsets.as<uint32_t>();
    uint32_t *counters = s.counters.as<uint32_t>();
    fill_u32_kernel<<<1024, 256, 0, stream>>>(status, status_words, 0u);
    fill_u32_kernel<<<1, 64, 0, stream>>>(counters, 64, 0u);
    digit_offsets_kernel<<<1, kSortRadix, 0, stream>>>(hist, offsets);
    uint32_t *kin = keys0, *vin = vals0, *kout = keys1, *vout = vals1;
    for (uint32_t p = 0; p < passes; ++p) {
        if (sort_items(n) == kSortItemsLarge)
            onesweep_pass_kernel<kSortItemsLarge><<<tiles, kSortBlock, 0, stream>>>(
                kin, vin, kout, vout, n, p * kSortRadixBits, offsets + p * kSortRadix,
                status + uint64_t(p) * tiles * kSortRadix, counters + p, err);
        else
            onesweep_pass_kernel<kSortItemsSmall><<<tiles, kSortBlock, 0, stream>>>(
                kin, vin, kout, vout, n, p * kSortRadixBits, offsets + p * kSortRadix,
                status + uint64_t(p) * tiles * kSortRadix, counters + p, err);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys_result = kin;  // an odd pass count leaves the result in buffer 1, an even one in buffer 0
    *vals_result = vin;
    return hipGetLastError();
}

}  // namespace pbrtm
