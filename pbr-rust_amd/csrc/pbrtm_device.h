// Device-side math and data layouts shared by the HLBVH build and traversal
// kernels (gfx950). Every FP32 operation here is written so that, compiled with
// -ffp-contract=off and correctly rounded divide/sqrt, it produces the same bits
// as the reference's host-side FP32 code (no fused multiply-add unless fmaf is
// written out, as the reference does in difference_of_products).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbrtm {

#define PBRTM_HD __host__ __device__ __forceinline__

constexpr uint32_t kMortonBits = 30;
constexpr uint32_t kTreeletShift = 9;     // 30-bit code >> 9 = 21-bit treelet id (128^3 grid)
constexpr uint32_t kNumBuckets = 24;      // top-level SAH buckets
constexpr uint32_t kSmallNode = 64;       // nodes up to this size finish inside one wave
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kNoHit = 0xFFFFFFFFu;

struct f3 {
    float x, y, z;
    PBRTM_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

PBRTM_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
PBRTM_HD f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PBRTM_HD f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PBRTM_HD f3 mul3(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
PBRTM_HD f3 abs3(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
PBRTM_HD f3 neg3(f3 a) { return {-a.x, -a.y, -a.z}; }
// std::min/std::max semantics: min(a, b) = (b < a) ? b : a.
PBRTM_HD float smin(float a, float b) { return (b < a) ? b : a; }
PBRTM_HD float smax(float a, float b) { return (a < b) ? b : a; }
PBRTM_HD f3 min3(f3 a, f3 b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
PBRTM_HD f3 max3(f3 a, f3 b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
PBRTM_HD float sel3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Exact FP32 a*b - c*d with one rounding error compensated (Kahan's trick).
PBRTM_HD float dop(float a, float b, float c, float d) {
    float cd = c * d;
    return fmaf(a, b, -cd) + fmaf(-c, d, cd);
}
PBRTM_HD float sop(float a, float b, float c, float d) {
    float cd = c * d;
    return fmaf(a, b, cd) + fmaf(c, d, -cd);
}
PBRTM_HD float dot3(f3 a, f3 b) { return fmaf(a.x, b.x, sop(a.y, b.y, a.z, b.z)); }
PBRTM_HD f3 cross3(f3 a, f3 b) {
    return {dop(a.y, b.z, a.z, b.y), dop(a.z, b.x, a.x, b.z), dop(a.x, b.y, a.y, b.x)};
}
PBRTM_HD float sqlen3(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// n * 2^-24 / (1 - n * 2^-24), rounded once each, as FP32 constants.
PBRTM_HD float gamma_n(int n) {
    const float eps = 5.9604644775390625e-08f;  // 2^-24
    return (float(n) * eps) / (1.0f - float(n) * eps);
}

// Axis-aligned box, empty = (+FLT_MAX, lowest).
struct Box {
    f3 lo, hi;
};
PBRTM_HD Box box_empty() {
    const float mx = 3.402823466e+38f;
    return {{mx, mx, mx}, {-mx, -mx, -mx}};
}
PBRTM_HD Box box_union(Box a, Box b) { return {min3(a.lo, b.lo), max3(a.hi, b.hi)}; }
PBRTM_HD Box box_add(Box a, f3 p) { return {min3(a.lo, p), max3(a.hi, p)}; }
PBRTM_HD f3 box_centroid(Box b) { return mul3(0.5f, add3(b.lo, b.hi)); }
PBRTM_HD int box_max_dim(Box b) {
    f3 d = sub3(b.hi, b.lo);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
}
// Bounds3::surface_area computes 2.0 * (FP32 sum) in double and rounds back; doubling
// is exact in both formats (an overflow gives +inf in both), so the FP32 product is
// the same value.
PBRTM_HD float box_area(Box b) {
    f3 d = sub3(b.hi, b.lo);
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}

// The SAH cost 0.125 + x (hlbvh.cu:711-731 adds the double 0.125 to an FP32 x and
// rounds to FP32): below 2^50 the double sum is exact, so its one rounding is the
// FP32 sum's; above, 0.125 is lost in both.
PBRTM_HD float sah_cost(float x) { return 0.125f + x; }
PBRTM_HD Box tri_box(f3 p0, f3 p1, f3 p2) {
    Box b{p0, p0};
    b.lo = min3(b.lo, p1); b.hi = max3(b.hi, p1);
    b.lo = min3(b.lo, p2); b.hi = max3(b.hi, p2);
    return b;
}

// FP32 -> u32 truncation that saturates (negative and NaN -> 0).
PBRTM_HD uint32_t f2u_sat(float f) {
    if (!(f > 0.f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return uint32_t(f);
}

// Spread the low 10 bits of x to bit positions 0,3,...,27; x == 1024 maps to 1023.
PBRTM_HD uint32_t spread3(uint32_t x) {
    x -= (x == 1024u);
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
PBRTM_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (spread3(z) << 2) | (spread3(y) << 1) | spread3(x);
}

// Order-preserving map FP32 <-> u32 so that min/max reductions can use integer atomics.
PBRTM_HD uint32_t f2ord(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
PBRTM_HD float ord2f(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __builtin_bit_cast(float, u);
}

// 32-byte element record used by every split pass: a box plus two words.
// Treelets: a = first primitive position, b = primitive count.
// Primitives: a = primitive id, b = 1.
struct alignas(16) Rec {
    float lo[3];
    uint32_t a;
    float hi[3];
    uint32_t b;
};
PBRTM_HD Box rec_box(const Rec &r) { return {{r.lo[0], r.lo[1], r.lo[2]}, {r.hi[0], r.hi[1], r.hi[2]}}; }

// Binary build node, 32 bytes. Interior: a = left id, b = right id | axis << 30.
// Leaf: a = kLeafFlag | first position, b = primitive count.
struct alignas(16) BNode {
    float lo[3];
    uint32_t a;
    float hi[3];
    uint32_t b;
};

// A node whose subtree is still to be built: its element range in one of the two
// record buffers, its preassigned node id, and its primitive count.
struct WorkItem {
    uint32_t begin, end;
    uint32_t node;
    uint32_t prims;
    uint32_t buf;   // 0 or 1: which record buffer holds [begin, end)
    uint32_t pad[3];
};

}  // namespace pbrtm
