// Ray/box and ray/triangle tests, hit-point geometry and the PCG32 stream used
// by the traversal and bounce kernels. FP32 operation order follows the
// reference so results are bit-identical under -ffp-contract=off.
#pragma once

#include "pbrtm_device.h"

namespace pbrtm {

// Far slab distances are widened as bounds3.h:159-180 does: t *= 1.0 + 2.0 * gamma(3),
// an FP64 product (the factor is a double) rounded back to FP32. With
// d = 2 * gamma(3), which is an FP32 number, one fused multiply-add t + t * d
// rounds the same exact product once: for every FP32 t (all 2^32 bit patterns,
// checked exhaustively, NaNs aside) fmaf(t, d, t) equals float(double(t) * (1 + d)).
__device__ __forceinline__ float widen_far(float t) { return fmaf(t, 2.0f * gamma_n(3), t); }

// Slab test of bounds3.h:146-195.
__device__ __forceinline__ bool box_hit(const float lo[3], const float hi[3], f3 o, f3 inv, const int neg[3],
                                        float ray_t_max) {
    float t_min = ((neg[0] ? hi[0] : lo[0]) - o.x) * inv.x;
    float t_max = ((neg[0] ? lo[0] : hi[0]) - o.x) * inv.x;
    float ty_min = ((neg[1] ? hi[1] : lo[1]) - o.y) * inv.y;
    float ty_max = ((neg[1] ? lo[1] : hi[1]) - o.y) * inv.y;
    t_max = widen_far(t_max);
    ty_max = widen_far(ty_max);
    if (t_min > ty_max || ty_min > t_max) return false;
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    float tz_min = ((neg[2] ? hi[2] : lo[2]) - o.z) * inv.z;
    float tz_max = ((neg[2] ? lo[2] : hi[2]) - o.z) * inv.z;
    tz_max = widen_far(tz_max);
    if (t_min > tz_max || tz_min > t_max) return false;
    if (tz_min > t_min) t_min = tz_min;
    if (tz_max < t_max) t_max = tz_max;
    return t_min < ray_t_max && t_max > 0.0f;
}

struct TriHit {
    float b0, b1, b2, t;
};

__device__ __forceinline__ float maxc(f3 a) { return smax(smax(a.x, a.y), a.z); }

// Watertight ray/triangle test of shapes/triangle.cu:217-328 (pbrt-v4).
// robust (nullable): set when the hit is clear of every box test a BVH traversal
// makes with this t_max: the ray is inside each edge by more than the edge
// functions' error bound deltaE, so the exact ray meets the triangle, and the hit
// lies more than the t error bound deltaT (and the slab distances' own rounding)
// before t_max, so every box holding the triangle has its entry distance below
// t_max (its far distances are widened, bounds3.h:159-180).
__device__ __forceinline__ bool tri_hit(f3 ro, f3 rd, float t_max, f3 p0, f3 p1, f3 p2, TriHit *out,
                                        bool *robust = nullptr) {
    if (sqlen3(cross3(sub3(p2, p0), sub3(p1, p0))) == 0.0f) return false;
    f3 a = sub3(p0, ro), b = sub3(p1, ro), c = sub3(p2, ro);
    const f3 ad = abs3(rd);
    const int kz = (ad.x > ad.y) ? ((ad.x > ad.z) ? 0 : 2) : ((ad.y > ad.z) ? 1 : 2);
    const int kx = kz == 2 ? 0 : kz + 1;
    const int ky = kx == 2 ? 0 : kx + 1;
    const f3 d = {sel3(rd, kx), sel3(rd, ky), sel3(rd, kz)};
    a = {sel3(a, kx), sel3(a, ky), sel3(a, kz)};
    b = {sel3(b, kx), sel3(b, ky), sel3(b, kz)};
    c = {sel3(c, kx), sel3(c, ky), sel3(c, kz)};
    const float Sx = -d.x / d.z, Sy = -d.y / d.z, Sz = 1 / d.z;
    a.x += Sx * a.z; a.y += Sy * a.z;
    b.x += Sx * b.z; b.y += Sy * b.z;
    c.x += Sx * c.z; c.y += Sy * c.z;
    float e0 = dop(b.x, c.y, b.y, c.x);
    float e1 = dop(c.x, a.y, c.y, a.x);
    float e2 = dop(a.x, b.y, a.y, b.x);
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {  // exact FP64 edge functions on the boundary
        e0 = float(double(c.y) * double(b.x) - double(c.x) * double(b.y));
        e1 = float(double(a.y) * double(c.x) - double(a.x) * double(c.y));
        e2 = float(double(b.y) * double(a.x) - double(b.x) * double(a.y));
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    const float det = e0 + e1 + e2;
    if (det == 0) return false;
    a.z *= Sz; b.z *= Sz; c.z *= Sz;
    const float tScaled = e0 * a.z + e1 * b.z + e2 * c.z;
    if (det < 0 && (tScaled >= 0 || tScaled < t_max * det)) return false;
    if (det > 0 && (tScaled <= 0 || tScaled > t_max * det)) return false;
    const float invDet = 1 / det;
    const float t = tScaled * invDet;
    const float maxZt = maxc(abs3(f3{a.z, b.z, c.z}));
    const float deltaZ = gamma_n(3) * maxZt;
    const float maxXt = maxc(abs3(f3{a.x, b.x, c.x}));
    const float maxYt = maxc(abs3(f3{a.y, b.y, c.y}));
    const float deltaX = gamma_n(5) * (maxXt + maxZt);
    const float deltaY = gamma_n(5) * (maxYt + maxZt);
    const float deltaE = 2 * (gamma_n(2) * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
    const float maxE = maxc(abs3(f3{e0, e1, e2}));
    const float deltaT = 3 * (gamma_n(3) * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * fabsf(invDet);
    if (t <= deltaT) return false;
    if (robust) {
        const float minE = fminf(fminf(fabsf(e0), fabsf(e1)), fabsf(e2));
        *robust = minE > deltaE && (t + deltaT) * (1.0f + 8.0f * gamma_n(3)) < t_max;
    }
    out->b0 = e0 * invDet;
    out->b1 = e1 * invDet;
    out->b2 = e2 * invDet;
    out->t = t;
    return true;
}

// ------------------------------------------------ directed rounding, offsets

PBRTM_HD float next_up(float v) {
    if (isinf(v) && v > 0.f) return v;
    if (v == -0.f) v = 0.f;
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    u = (v >= 0.f) ? u + 1 : u - 1;
    return __builtin_bit_cast(float, u);
}
PBRTM_HD float next_down(float v) {
    if (isinf(v) && v < 0.f) return v;
    if (v == 0.f) v = -0.f;
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    u = (v > 0.f) ? u - 1 : u + 1;
    return __builtin_bit_cast(float, u);
}
// Exact a+b rounded toward +inf / -inf: RN(a+b) corrected by the sign of the
// TwoSum remainder.
PBRTM_HD float add_up(float a, float b) {
    const float s = a + b;
    if (!isfinite(s)) return s;
    const float bb = s - a;
    const float r = (a - (s - bb)) + (b - bb);
    return r > 0.f ? next_up(s) : s;
}
PBRTM_HD float add_down(float a, float b) {
    const float s = a + b;
    if (!isfinite(s)) return s;
    const float bb = s - a;
    const float r = (a - (s - bb)) + (b - bb);
    return r < 0.f ? next_down(s) : s;
}

struct Ival {
    float lo, hi;
};
PBRTM_HD Ival ival(float v, float err) {
    if (err == 0.f) return {v, v};
    return {add_down(v, -err), add_up(v, err)};
}
PBRTM_HD float ival_mid(Ival i) { return (i.lo + i.hi) / 2; }
PBRTM_HD float ival_half(Ival i) { return (i.hi - i.lo) / 2; }

// Hit point as an interval box (triangle.cu:376-384) and the geometric normal
// normalize((p0 - p2) x (p1 - p2)) (triangle.cu:391).
struct HitGeom {
    Ival x, y, z;
    f3 n;
};
PBRTM_HD HitGeom hit_geom(const TriHit &h, f3 p0, f3 p1, f3 p2) {
    const f3 ph = add3(add3(mul3(h.b0, p0), mul3(h.b1, p1)), mul3(h.b2, p2));
    const f3 pa = add3(add3(abs3(mul3(h.b0, p0)), abs3(mul3(h.b1, p1))), abs3(mul3(h.b2, p2)));
    const f3 pe = mul3(gamma_n(7), pa);
    const f3 nn = cross3(sub3(p0, p2), sub3(p1, p2));
    const float len = sqrtf(sqlen3(nn));
    return {ival(ph.x, pe.x), ival(ph.y, pe.y), ival(ph.z, pe.z), f3{nn.x / len, nn.y / len, nn.z / len}};
}

// Ray origin pushed past the error box along n (ray.h:22-43).
PBRTM_HD f3 offset_origin(const HitGeom &g, f3 n, f3 w) {
    const f3 err = {ival_half(g.x), ival_half(g.y), ival_half(g.z)};
    const float dd = fabsf(dot3(n, err));
    f3 off = mul3(dd, n);
    if (dot3(n, w) < 0) off = neg3(off);
    f3 po = add3(f3{ival_mid(g.x), ival_mid(g.y), ival_mid(g.z)}, off);
    if (off.x > 0) po.x = next_up(po.x); else if (off.x < 0) po.x = next_down(po.x);
    if (off.y > 0) po.y = next_up(po.y); else if (off.y < 0) po.y = next_down(po.y);
    if (off.z > 0) po.z = next_up(po.z); else if (off.z < 0) po.z = next_down(po.z);
    return po;
}

// Orthonormal basis around v (vector3.h:160-166).
PBRTM_HD void basis(f3 v, f3 *a, f3 *b) {
    const float sign = copysignf(1.0f, v.z);
    const float aa = -1.0f / (sign + v.z);
    const float bb = v.x * v.y * aa;
    *a = {1 + sign * (v.x * v.x) * aa, sign * bb, -sign * v.x};
    *b = {bb, sign + (v.y * v.y) * aa, -v.y};
}

// ------------------------------------------------------------------ PCG32

struct Pcg {
    uint64_t state, inc;
    PBRTM_HD uint32_t next() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    PBRTM_HD float uniform() {
        const float f = float(next()) * 0x1p-32f;
        return f < 0x1.fffffep-1f ? f : 0x1.fffffep-1f;
    }
};
PBRTM_HD uint64_t mix64(uint64_t v) {
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}
PBRTM_HD Pcg pcg_seq(uint64_t seq) {
    Pcg r{0u, (seq << 1u) | 1u};
    r.next();
    r.state += mix64(seq);
    r.next();
    return r;
}

}  // namespace pbrtm
