// TEST INFRASTRUCTURE ONLY — part of the CPU oracle.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// the oracle, and only as the checker. The product library never links it.
//
// Scalar FP32 helpers that restate the reference's float semantics exactly.
// Built with -O2 -ffp-contract=off -fno-fast-math: every a*b+c below is two
// roundings unless std::fma is written out, as in the reference host code.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <algorithm>

namespace orc {

struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

// util/basic_math.h:33-35 — gamma(n) evaluated in FP32, MachineEpsilon = 2^-24.
inline float gammaf(int n) {
    const float eps = std::numeric_limits<float>::epsilon() * 0.5;
    return (float(n) * eps) / (1.0f - float(n) * eps);
}

// util/basic_math.h:124-138 — difference/sum of products with explicit FMA.
inline float dop(float a, float b, float c, float d) {
    float cd = c * d;
    float diff = std::fma(a, b, -cd);
    float err = std::fma(-c, d, cd);
    return diff + err;
}
inline float sop(float a, float b, float c, float d) {
    float cd = c * d;
    float s = std::fma(a, b, cd);
    float err = std::fma(c, d, -cd);
    return s + err;
}

// euclidean_space/vector3.h:140-154 — dot is FMA(x, x', SoP(y, y', z, z')).
inline float dot(V3 a, V3 b) { return std::fma(a.x, b.x, sop(a.y, b.y, a.z, b.z)); }
inline V3 cross(V3 a, V3 b) {
    return {dop(a.y, b.z, a.z, b.y), dop(a.z, b.x, a.x, b.z), dop(a.x, b.y, a.y, b.x)};
}
// vector3.h:122-132 — squared_length is the plain x*x + y*y + z*z.
inline float sqlen(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 normalize(V3 a) { return divs(a, std::sqrt(sqlen(a))); }
inline int max_component_index(V3 a) {  // vector3.h:102-104
    return (a.x > a.y) ? ((a.x > a.z) ? 0 : 2) : ((a.y > a.z) ? 1 : 2);
}
inline float max_component_value(V3 a) { return std::max(std::max(a.x, a.y), a.z); }

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// util/bit_arithmetic.h:55-114 — one-ulp steps.
inline float next_float_up(float v) {
    if (std::isinf(v) && v > 0.f) return v;
    if (v == -0.f) v = 0.f;
    uint32_t ui = f2u(v);
    if (v >= 0.f) ++ui; else --ui;
    return u2f(ui);
}
inline float next_float_down(float v) {
    if (std::isinf(v) && v < 0.f) return v;
    if (v == 0.f) v = -0.f;
    uint32_t ui = f2u(v);
    if (v > 0.f) --ui; else ++ui;
    return u2f(ui);
}

// util/bit_arithmetic.h:137-171. The reference runs these on the device as
// __fadd_ru / __fadd_rd: the exact sum rounded toward +inf / -inf. That is
// restated here with the TwoSum error term (Knuth): s = RN(a+b) is already the
// directed result unless the exact remainder points the other way.
inline float add_round_up(float a, float b) {
    float s = a + b;
    if (!std::isfinite(s)) return s;
    float bb = s - a;
    float err = (a - (s - bb)) + (b - bb);
    return err > 0.f ? next_float_up(s) : s;
}
inline float add_round_down(float a, float b) {
    float s = a + b;
    if (!std::isfinite(s)) return s;
    float bb = s - a;
    float err = (a - (s - bb)) + (b - bb);
    return err < 0.f ? next_float_down(s) : s;
}

// util/interval.h:23-35 and euclidean_space/point3fi.h:32-48.
struct Interval {
    float lo, hi;
};
inline Interval from_value_and_error(float v, float err) {
    if (err == 0.f) return {v, v};
    return {add_round_down(v, -err), add_round_up(v, err)};
}
inline float midpoint(Interval i) { return (i.lo + i.hi) / 2; }
inline float half_width(Interval i) { return (i.hi - i.lo) / 2; }

struct P3fi {
    Interval x, y, z;
};
inline P3fi make_p3fi(V3 p, V3 e) {
    return {from_value_and_error(p.x, e.x), from_value_and_error(p.y, e.y),
            from_value_and_error(p.z, e.z)};
}
inline V3 p3fi_point(const P3fi &p) { return {midpoint(p.x), midpoint(p.y), midpoint(p.z)}; }
inline V3 p3fi_error(const P3fi &p) { return {half_width(p.x), half_width(p.y), half_width(p.z)}; }

// base/ray.h:22-43 — push the origin off the surface along n past the error box.
inline V3 offset_ray_origin(const P3fi &pi, V3 n, V3 w) {
    float d = std::fabs(dot(n, p3fi_error(pi)));
    V3 offset = scale(d, n);
    if (dot(n, w) < 0) offset = neg(offset);
    V3 po = add(p3fi_point(pi), offset);
    for (int i = 0; i < 3; ++i) {
        if (offset[i] > 0) po[i] = next_float_up(po[i]);
        else if (offset[i] < 0) po[i] = next_float_down(po[i]);
    }
    return po;
}

// Bounds3 (euclidean_space/bounds3.h). Default = empty: p_min = +max, p_max = lowest.
struct B3 {
    V3 lo, hi;
};
inline B3 b3_empty() {
    const float mx = std::numeric_limits<float>::max();
    const float mn = std::numeric_limits<float>::lowest();
    return {{mx, mx, mx}, {mn, mn, mn}};
}
inline B3 b3_union(B3 a, B3 b) { return {vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
inline B3 b3_add_point(B3 a, V3 p) { return {vmin(a.lo, p), vmax(a.hi, p)}; }
// bounds3.h:101-103 — 0.5 * (p_min + p_max) with 0.5 converted to FP32.
inline V3 b3_centroid(B3 b) { return scale(0.5f, add(b.lo, b.hi)); }
// bounds3.h:123-132
inline int b3_max_dimension(B3 b) {
    V3 d = sub(b.hi, b.lo);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
}
// bounds3.h:53-57 — FP32 products and sums, then 2.0 * (double) rounded back.
inline float b3_surface_area(B3 b) {
    V3 d = sub(b.hi, b.lo);
    return float(2.0 * (d.x * d.y + d.x * d.z + d.y * d.z));
}
// bounds3.h:33-46 — bounds of a triangle's three points.
inline B3 b3_of_tri(V3 p0, V3 p1, V3 p2) {
    B3 b{p0, p0};
    b.lo = vmin(b.lo, p1); b.hi = vmax(b.hi, p1);
    b.lo = vmin(b.lo, p2); b.hi = vmax(b.hi, p2);
    return b;
}
// bounds3.h:105-121
inline V3 b3_offset(B3 b, V3 p) {
    V3 o = sub(p, b.lo);
    if (b.hi.x > b.lo.x) o.x /= (b.hi.x - b.lo.x);
    if (b.hi.y > b.lo.y) o.y /= (b.hi.y - b.lo.y);
    if (b.hi.z > b.lo.z) o.z /= (b.hi.z - b.lo.z);
    return o;
}

// bounds3.h:146-195 — slab test, far planes widened by (1 + 2*gamma(3)) in FP64.
inline bool b3_fast_intersect(const B3 &b, V3 o, float ray_t_max, V3 inv_dir, const int neg[3]) {
    const double widen = 1.0 + 2.0 * double(gammaf(3));
    const V3 *c[2] = {&b.lo, &b.hi};
    float t_min = ((*c[neg[0]]).x - o.x) * inv_dir.x;
    float t_max = ((*c[1 - neg[0]]).x - o.x) * inv_dir.x;
    float ty_min = ((*c[neg[1]]).y - o.y) * inv_dir.y;
    float ty_max = ((*c[1 - neg[1]]).y - o.y) * inv_dir.y;
    t_max = float(double(t_max) * widen);
    ty_max = float(double(ty_max) * widen);
    if (t_min > ty_max || ty_min > t_max) return false;
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    float tz_min = ((*c[neg[2]]).z - o.z) * inv_dir.z;
    float tz_max = ((*c[1 - neg[2]]).z - o.z) * inv_dir.z;
    tz_max = float(double(tz_max) * widen);
    if (t_min > tz_max || tz_min > t_max) return false;
    if (tz_min > t_min) t_min = tz_min;
    if (tz_max < t_max) t_max = tz_max;
    return t_min < ray_t_max && t_max > 0.0;
}

// Saturating FP32 -> u32 truncation (what the reference's device conversion does:
// negative and NaN give 0, values >= 2^32 give 0xFFFFFFFF).
inline uint32_t f2u_sat(float f) {
    if (!(f > 0.f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return uint32_t(f);
}

// util/basic_math.h:140-165 — spread 10 bits to every third bit; 1024 clamps to 1023.
inline uint32_t left_shift3(uint32_t x) {
    if (x == (1u << 10)) --x;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
inline uint32_t encode_morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (left_shift3(z) << 2) | (left_shift3(y) << 1) | left_shift3(x);
}

}  // namespace orc
