// Host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

namespace pbrtm {

// Device error bits, OR-ed into a device word by kernels and checked by the host.
enum DeviceError : uint32_t {
    kErrNonFinite = 1u << 0,        // a vertex coordinate is Inf or NaN
    kErrDegenerateCentroids = 1u << 1,  // >1 treelet with one centroid on the split axis
    kErrSpin = 1u << 2,             // a bounded look-back spin ran out
    kErrStack = 1u << 3,            // traversal stack overflow
    kErrSplit = 1u << 4,            // SAH produced an empty side (unreachable for finite input)
    kErrCapacity = 1u << 5,         // a work list overflowed its capacity
};

std::string device_error_text(uint32_t bits);

// Growable device buffer owned by a context or a BVH handle.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need);
    void release();
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
};

// Exclusive scan of n uint32 values on `stream`. `total` (device, may be null)
// receives the sum. Uses `scratch` for block sums.
hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total,
                              DevBuf &scratch, hipStream_t stream);

// Stable LSD radix sort of 30-bit keys with u32 values, three 10-bit onesweep
// passes. `hist` must hold the three digit histograms (3 x 1024, counted by the
// caller); keys/vals ping-pong between the two buffers; *keys_result/*vals_result
// name the buffers holding the sorted output.
struct SortScratch {
    DevBuf status, offsets, counters;
};
hipError_t onesweep_sort_30(uint32_t *keys0, uint32_t *vals0, uint32_t *keys1, uint32_t *vals1,
                            uint32_t n, const uint32_t *hist, uint32_t *err, SortScratch &s,
                            hipStream_t stream, uint32_t **keys_result, uint32_t **vals_result);
// Sizes the sort's scratch for n keys (so that later calls allocate nothing).
hipError_t onesweep_reserve(SortScratch &s, uint32_t n);
// The same over the low 10 * passes bits (passes <= 3); hist holds all three digit histograms.
hipError_t onesweep_sort(uint32_t *keys0, uint32_t *vals0, uint32_t *keys1, uint32_t *vals1, uint32_t n,
                         const uint32_t *hist, uint32_t *err, SortScratch &s, hipStream_t stream,
                         uint32_t **keys_result, uint32_t **vals_result, uint32_t passes);
constexpr uint32_t kSortRadixBits = 10;
constexpr uint32_t kSortRadix = 1u << kSortRadixBits;
constexpr uint32_t kSortPasses = 3;

}  // namespace pbrtm
