// Device-wide exclusive scan of uint32 (reduce, spine, down-sweep). Used for
// treelet numbering and chunk offsets; HBM-bound: 4 B read twice + 4 B written.
#include "pbrtm_internal.h"
#include "wave_ops.h"

namespace pbrtm {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const uint32_t *__restrict__ in, uint32_t n,
                                                                uint32_t *__restrict__ block_sums) {
    __shared__ uint32_t s_waves[kScanBlock / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t i = base + uint64_t(k) * kScanBlock + threadIdx.x;
        if (i < n) sum += in[i];
    }
    sum = wave_sum(sum);
    if (__lane_id() == 0) s_waves[threadIdx.x / kWave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kScanBlock / kWave; ++i) t += s_waves[i];
        block_sums[blockIdx.x] = t;
    }
}

// One workgroup scans the block sums in place (any count), writes the total.
__global__ __launch_bounds__(1024) void scan_spine_kernel(uint32_t *__restrict__ sums, uint32_t nb,
                                                          uint32_t *__restrict__ total) {
    __shared__ uint32_t s_waves[1024 / kWave];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_add<1024>(v, s_waves, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_down_kernel(const uint32_t *in,
                                                               uint32_t *out, uint32_t n,
                                                               const uint32_t *__restrict__ block_offsets) {
    __shared__ uint32_t s_tile[kScanTile + kScanTile / 32];
    __shared__ uint32_t s_waves[kScanBlock / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    auto pad = [](uint32_t i) { return i + (i >> 5); };
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {  // striped, coalesced load into LDS
        uint32_t j = k * kScanBlock + threadIdx.x;
        uint64_t i = base + j;
        s_tile[pad(j)] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {  // blocked: thread owns 16 consecutive items
        v[k] = s_tile[pad(threadIdx.x * kScanItems + k)];
        uint32_t t = v[k];
        v[k] = run;
        run += t;
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_add<kScanBlock>(run, s_waves, &tot) + block_offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s_tile[pad(threadIdx.x * kScanItems + k)] = v[k] + ex;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t j = k * kScanBlock + threadIdx.x;
        uint64_t i = base + j;
        if (i < n) out[i] = s_tile[pad(j)];
    }
}

// Small inputs: one workgroup walks the tiles in order with a running carry
// (one launch instead of three; in place is fine, a tile is read before it is
// written).
__global__ __launch_bounds__(kScanBlock) void scan_single_kernel(const uint32_t *in, uint32_t *out, uint32_t n,
                                                                 uint32_t *__restrict__ total) {
    __shared__ uint32_t s_tile[kScanTile + kScanTile / 32];
    __shared__ uint32_t s_waves[kScanBlock / kWave];
    auto pad = [](uint32_t i) { return i + (i >> 5); };
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanTile) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t j = k * kScanBlock + threadIdx.x;
            s_tile[pad(j)] = base + j < n ? in[base + j] : 0u;
        }
        __syncthreads();
        uint32_t v[kScanItems];
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t t = s_tile[pad(threadIdx.x * kScanItems + k)];
            v[k] = run;
            run += t;
        }
        uint32_t tot;
        const uint32_t ex = block_exclusive_add<kScanBlock>(run, s_waves, &tot) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) s_tile[pad(threadIdx.x * kScanItems + k)] = v[k] + ex;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t j = k * kScanBlock + threadIdx.x;
            if (base + j < n) out[base + j] = s_tile[pad(j)];
        }
        carry += tot;
        __syncthreads();  // s_tile and s_waves are reused by the next tile
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

constexpr uint32_t kScanSingleMax = 16 * kScanTile;

}  // namespace

hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total,
                              DevBuf &scratch, hipStream_t stream) {
    if (n <= kScanSingleMax) {
        scan_single_kernel<<<1, kScanBlock, 0, stream>>>(in, out, n, total);
        return hipGetLastError();
    }
    const uint32_t nb = n == 0 ? 0 : (n + kScanTile - 1) / kScanTile;
    hipError_t e = scratch.reserve(size_t(nb + 1) * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    uint32_t *sums = scratch.as<uint32_t>();
    if (nb > 0) scan_reduce_kernel<<<nb, kScanBlock, 0, stream>>>(in, n, sums);
    scan_spine_kernel<<<1, 1024, 0, stream>>>(sums, nb, total);
    if (nb > 0) scan_down_kernel<<<nb, kScanBlock, 0, stream>>>(in, out, n, sums);
    return hipGetLastError();
}

}  // namespace pbrtm
