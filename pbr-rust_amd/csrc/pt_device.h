// Device-side pieces of the wavefront path tracer: 4-wavelength spectra,
// piecewise-linear spectrum lookup, the stratified sampler, Mitchell filter
// sampling, diffuse and specular-conductor BSDFs, spherical-triangle light
// sampling and the MIS helpers. FP32 operation order and double promotions
// follow the reference (cited per function) so that, built with
// -ffp-contract=off, every result is bit-identical to the host restatement.
// Transcendentals are evaluated in double precision and rounded once to float.
#pragma once

#include "pbrtm_device.h"
#include "pbrtm_geometry.h"

namespace pbrtm {
namespace pt {

constexpr int kNS = 4;  // wavelengths per path
constexpr float kLambdaMin = 360.f, kLambdaMax = 830.f;
constexpr float kOneMinusEps = 0x1.fffffep-1f;
constexpr float kMinSolidAngle = 3e-4f, kMaxSolidAngle = 6.22f;

// acos(-1) rounded to float, and float(1.0 / double(pi_f)).
constexpr float kPi = 3.14159274101257324f;
constexpr float kInvPi = float(1.0 / double(kPi));

PBRTM_HD float tsin(float x) { return float(sin(double(x))); }
PBRTM_HD float tcos(float x) { return float(cos(double(x))); }
PBRTM_HD float tasin(float x) { return float(asin(double(x))); }
PBRTM_HD float tatan2(float y, float x) { return float(atan2(double(y), double(x))); }
PBRTM_HD float tatanh(float x) { return float(atanh(double(x))); }
PBRTM_HD float tcosh(float x) { return float(cosh(double(x))); }

PBRTM_HD float sqrf(float v) { return v * v; }
PBRTM_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
PBRTM_HD float lerpf(float t, float a, float b) { return (1 - t) * a + t * b; }
PBRTM_HD float safe_sqrtf(float x) { return sqrtf(smax(0.0f, x)); }
PBRTM_HD float len3(f3 a) { return sqrtf(sqlen3(a)); }
PBRTM_HD f3 norm3(f3 a) {
    const float l = len3(a);
    return {a.x / l, a.y / l, a.z / l};
}
PBRTM_HD float absdot3(f3 a, f3 b) { return fabsf(dot3(a, b)); }
PBRTM_HD bool zero3(f3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }

// vector3.h:233-240 and :243-245
PBRTM_HD float angle_between(f3 v1, f3 v2) {
    if (dot3(v1, v2) < 0) return kPi - 2 * tasin(clampf(len3(add3(v1, v2)) / 2, -1, 1));
    return 2 * tasin(clampf(len3(sub3(v2, v1)) / 2, -1, 1));
}
PBRTM_HD f3 gram_schmidt(f3 v, f3 w) { return sub3(v, mul3(dot3(v, w), w)); }

// ------------------------------------------------------------ sampled spectra
struct S4 {
    float v[kNS];
};
PBRTM_HD S4 s4(float a) { return {{a, a, a, a}}; }
PBRTM_HD S4 s4mul(const S4 &a, const S4 &b) { return {{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}}; }
PBRTM_HD S4 s4scale(const S4 &a, float s) { return {{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }
PBRTM_HD S4 s4div(const S4 &a, float s) { return {{a.v[0] / s, a.v[1] / s, a.v[2] / s, a.v[3] / s}}; }
PBRTM_HD S4 s4add(const S4 &a, const S4 &b) { return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}}; }
PBRTM_HD bool s4pos(const S4 &a) { return a.v[0] > 0 || a.v[1] > 0 || a.v[2] > 0 || a.v[3] > 0; }
PBRTM_HD float s4max(const S4 &a) { return smax(smax(smax(a.v[0], a.v[1]), a.v[2]), a.v[3]); }
PBRTM_HD float s4avg(const S4 &a) { return (((0.0f + a.v[0]) + a.v[1]) + a.v[2] + a.v[3]) / kNS; }
PBRTM_HD S4 s4safediv(const S4 &a, const S4 &b) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = b.v[i] == 0.0f ? 0.0f : a.v[i] / b.v[i];
    return r;
}
PBRTM_HD float4 s4f(const S4 &a) { return make_float4(a.v[0], a.v[1], a.v[2], a.v[3]); }
PBRTM_HD S4 f4s(float4 a) { return {{a.x, a.y, a.z, a.w}}; }

// sampling.h:240-253 and sampled_wavelengths.h:25-42
PBRTM_HD void sample_wavelengths(float u, S4 *lam, S4 *pdf) {
    for (int i = 0; i < kNS; ++i) {
        float up = u + float(i) / kNS;
        if (up > 1) up -= 1;
        const float l = 538 - 138.888889f * tatanh(0.85691062f - 1.82750197f * up);
        lam->v[i] = l;
        pdf->v[i] = (l < kLambdaMin || l > kLambdaMax) ? 0.0f : 0.0039398042f / sqrf(tcosh(0.0072f * (l - 538)));
    }
}

// Sigmoid polynomial s(c0 l^2 + c1 l + c2) (rgb_sigmoid_polynomial.h), Horner with fma.
PBRTM_HD float sigmoid_eval(float c0, float c1, float c2, float l) {
    const float x = fmaf(l, fmaf(l, c0, c1), c2);
    if (isinf(x)) return x > 0 ? 1.f : 0.f;
    return .5f + x / (2 * sqrtf(1 + x * x));
}

// Piecewise-linear spectrum (piecewise_linear_spectrum.cu:92-101): the largest
// sample at or below l by the reference's binary search, then a lerp.
PBRTM_HD size_t find_interval_le(const float *lam, size_t sz, float l) {
    long size = (long)sz - 2, first = 1;
    while (size > 0) {
        const long half = size >> 1, middle = first + half;
        const bool r = lam[middle] <= l;
        first = r ? middle + 1 : first;
        size = r ? size - (half + 1) : half;
    }
    long v = first - 1;
    v = v < 0 ? 0 : v;
    v = v > (long)sz - 2 ? (long)sz - 2 : v;
    return (size_t)v;
}
// find_interval_le for a grid whose entries from index a on are evenly spaced
// (1 / step = inv_step): the guessed interval is checked against the grid, so the
// result is the binary search's in every case.
PBRTM_HD size_t find_interval_guess(const float *lam, size_t n, uint32_t a, float inv_step, float l) {
    const float la = lam[a];
    if (inv_step > 0.f && l >= la) {
        long g = long(a) + long((l - la) * inv_step);
        g = g > long(n) - 2 ? long(n) - 2 : g;
        if (lam[g] <= l && (g == long(n) - 2 || lam[g + 1] > l)) return size_t(g);
    }
    return find_interval_le(lam, n, l);
}
PBRTM_HD float pl_eval(const float *lam, const float *val, uint32_t n, float l) {
    if (n == 0 || l < lam[0] || l > lam[n - 1]) return 0;
    const size_t o = find_interval_le(lam, n, l);
    const float t = (l - lam[o]) / (lam[o + 1] - lam[o]);
    return lerpf(t, val[o], val[o + 1]);
}

// ------------------------------------------------------------ hashing and sampler
// MurmurHash64A of 4 or 8 bytes (util/hash.h), seed 0.
PBRTM_HD uint64_t murmur_tail(uint64_t h) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    h ^= h >> 47;
    h *= m;
    h ^= h >> 47;
    return h;
}
PBRTM_HD uint64_t hash_u32(uint32_t a) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    uint64_t h = 0ull ^ (4ull * m);
    h ^= uint64_t(a);
    h *= m;
    return murmur_tail(h);
}
PBRTM_HD uint64_t hash_u32x2(uint32_t a, uint32_t b) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    uint64_t h = 0ull ^ (8ull * m);
    uint64_t k = uint64_t(a) | (uint64_t(b) << 32);
    k *= m;
    k ^= k >> 47;
    k *= m;
    h ^= k;
    h *= m;
    return murmur_tail(h);
}
// Kensler's hashed permutation (util/permutation.h).
PBRTM_HD uint32_t permutation_element(uint32_t i, uint32_t l, uint32_t p) {
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893d;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3f;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1 | p >> 27;
        i *= 0x6935fa69;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3;
        i ^= (i & w) >> 2;
        i *= 0xc860a3df;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}
// PCG32 jump by delta steps (rng.h:135-148).
PBRTM_HD void pcg_advance(Pcg &r, uint64_t delta) {
    uint64_t cur_mult = 0x5851f42d4c957f2dULL, cur_plus = r.inc, acc_mult = 1u, acc_plus = 0u;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta /= 2;
    }
    r.state = acc_mult * r.state + acc_plus;
}

// Stratified sampler (samplers/stratified.h); the state that travels with a
// path is the PCG32 state and the dimension; the stream is re-derived from the pixel.
struct Sampler {
    Pcg rng;
    uint32_t pixel, sample, dim, spd;
    PBRTM_HD float get_1d() {
        const uint32_t spp = spd * spd;
        const uint32_t stratum = permutation_element(sample, spp, uint32_t(hash_u32x2(pixel, dim)));
        dim += 1;
        const float delta = rng.uniform();
        return (float(int(stratum)) + delta) / float(spp);
    }
    PBRTM_HD void get_2d(float *u0, float *u1) {
        const uint32_t spp = spd * spd;
        const uint32_t stratum = permutation_element(sample, spp, uint32_t(hash_u32x2(pixel, dim)));
        dim += 2;
        const uint32_t x = stratum % spd, y = stratum / spd;
        const float dx = rng.uniform();
        const float dy = rng.uniform();
        *u0 = (float(x) + dx) / float(spd);
        *u1 = (float(y) + dy) / float(spd);
    }
};
PBRTM_HD uint64_t sampler_inc(uint32_t pixel) { return (hash_u32(pixel) << 1u) | 1u; }

// ------------------------------------------------------------ power heuristic, bilinear warp
PBRTM_HD float power_heuristic(float f, float g) {
    const float ff = 1 * f, gg = 1 * g;
    if (isinf(sqrf(ff))) return 1;
    return sqrf(ff) / (sqrf(ff) + sqrf(gg));
}
PBRTM_HD float sample_linear(float u, float a, float b) {
    if (u == 0 && a == 0) return 0;
    const float x = u * (a + b) / (a + sqrtf(lerpf(u, sqrf(a), sqrf(b))));
    return smin(x, kOneMinusEps);
}
PBRTM_HD float bilinear_pdf(float p0, float p1, const float w[4]) {
    if (p0 < 0 || p0 > 1 || p1 < 0 || p1 > 1) return 0;
    if (w[0] + w[1] + w[2] + w[3] == 0) return 1;
    const float s = (1 - p0) * (1 - p1) * w[0] + p0 * (1 - p1) * w[1] + (1 - p0) * p1 * w[2] + p0 * p1 * w[3];
    return float(4.0 * double(s) / double(w[0] + w[1] + w[2] + w[3]));
}

// ------------------------------------------------------------ complex Fresnel
struct Cx {
    float re, im;
};
PBRTM_HD Cx cxadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
PBRTM_HD Cx cxsub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
PBRTM_HD Cx cxmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
PBRTM_HD Cx cxdiv(Cx a, Cx b) {
    const float s = 1 / (b.re * b.re + b.im * b.im);
    return {s * (a.re * b.re + a.im * b.im), s * (a.im * b.re - a.re * b.im)};
}
PBRTM_HD float cxnorm(Cx a) { return a.re * a.re + a.im * a.im; }
PBRTM_HD Cx cxsqrt(Cx z) {  // util/complex.h
    const float n = sqrtf(cxnorm(z));
    const float t1 = sqrtf(.5f * (n + fabsf(z.re)));
    const float t2 = .5f * z.im / t1;
    if (n == 0) return {0.f, 0.f};
    if (z.re >= 0) return {t1, t2};
    return {fabsf(t2), copysignf(t1, z.im)};
}
// util/scattering.h:56-72
PBRTM_HD float fr_complex(float cos_i, float eta_re, float eta_im) {
    const Cx eta{eta_re, eta_im};
    cos_i = clampf(cos_i, 0, 1);
    const float sin2_i = 1 - sqrf(cos_i);
    const Cx sin2_t = cxdiv(Cx{sin2_i, 0.f}, cxmul(eta, eta));
    const Cx cos_t = cxsqrt(cxsub(Cx{1.f, 0.f}, sin2_t));
    const Cx ci{cos_i, 0.f};
    const Cx r_parl = cxdiv(cxsub(cxmul(eta, ci), cos_t), cxadd(cxmul(eta, ci), cos_t));
    const Cx r_perp = cxdiv(cxsub(ci, cxmul(eta, cos_t)), cxadd(ci, cxmul(eta, cos_t)));
    return (cxnorm(r_parl) + cxnorm(r_perp)) / 2;
}

// ------------------------------------------------------------ diffuse disk / hemisphere
// sampling.h:61-83 and :219-224
PBRTM_HD f3 cosine_hemisphere(float u0, float u1) {
    const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    float dx = 0.f, dy = 0.f;
    if (!(ox == 0.0f && oy == 0.0f)) {
        const float pi4 = float(double(kPi) / 4.0);
        float r, theta;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            theta = pi4 * (oy / ox);
        } else {
            r = oy;
            const float pi2 = float(double(kPi) / 2.0);
            theta = pi2 - pi4 * (ox / oy);
        }
        dx = r * tcos(theta);
        dy = r * tsin(theta);
    }
    const float zz = float(1.0 - double(dx * dx) - double(dy * dy));
    return {dx, dy, sqrtf(smax(0.0f, zz))};
}

// util/sampling.h:203-216 (UniformInfiniteLight::sample_li)
PBRTM_HD f3 uniform_sphere(float u0, float u1) {
    const float z = 1 - 2 * u0;
    const float r = sqrtf(smax(0.0f, 1 - z * z));
    const float phi = 2 * kPi * u1;
    return {r * tcos(phi), r * tsin(phi), z};
}
constexpr float kUniformSpherePdf = float(1.0 / (4.0 * double(kPi)));

// ------------------------------------------------------------ spherical triangles
// shapes/triangle.h:25-27
PBRTM_HD float spherical_triangle_area(f3 a, f3 b, f3 c) {
    return float(fabs(2.0 * double(tatan2(dot3(a, cross3(b, c)), 1 + dot3(a, b) + dot3(a, c) + dot3(b, c)))));
}
// util/sampling.h:109-201
PBRTM_HD void sample_spherical_triangle(float out[3], float *pdf, const f3 v[3], f3 p, float u0, float u1) {
    *pdf = 0;
    const f3 a = norm3(sub3(v[0], p)), b = norm3(sub3(v[1], p)), c = norm3(sub3(v[2], p));
    f3 n_ab = cross3(a, b), n_bc = cross3(b, c), n_ca = cross3(c, a);
    if (sqlen3(n_ab) == 0 || sqlen3(n_bc) == 0 || sqlen3(n_ca) == 0) return;
    n_ab = norm3(n_ab);
    n_bc = norm3(n_bc);
    n_ca = norm3(n_ca);
    const float alpha = angle_between(n_ab, neg3(n_ca));
    const float beta = angle_between(n_bc, neg3(n_ab));
    const float gamma = angle_between(n_ca, neg3(n_bc));
    const float A_pi = alpha + beta + gamma;
    const float Ap_pi = lerpf(u0, kPi, A_pi);
    const float A = A_pi - kPi;
    *pdf = (A <= 0) ? 0 : 1 / A;
    const float cosAlpha = tcos(alpha), sinAlpha = tsin(alpha);
    const float sinPhi = tsin(Ap_pi) * cosAlpha - tcos(Ap_pi) * sinAlpha;
    const float cosPhi = tcos(Ap_pi) * cosAlpha + tsin(Ap_pi) * sinAlpha;
    const float k1 = cosPhi + cosAlpha;
    const float k2 = sinPhi - sinAlpha * dot3(a, b);
    float cosBp = (k2 + (dop(k2, cosPhi, k1, sinPhi)) * cosAlpha) / ((sop(k2, sinPhi, k1, cosPhi)) * sinAlpha);
    cosBp = clampf(cosBp, -1, 1);
    const float sinBp = safe_sqrtf(1 - sqrf(cosBp));
    const f3 cp = add3(mul3(cosBp, a), mul3(sinBp, norm3(gram_schmidt(c, a))));
    const float cosTheta = 1 - u1 * (1 - dot3(cp, b));
    const float sinTheta = safe_sqrtf(1 - sqrf(cosTheta));
    const f3 w = add3(mul3(cosTheta, b), mul3(sinTheta, norm3(gram_schmidt(cp, b))));
    const f3 e1 = sub3(v[1], v[0]), e2 = sub3(v[2], v[0]);
    const f3 s1 = cross3(w, e2);
    const float divisor = dot3(s1, e1);
    if (divisor == 0) {
        out[0] = out[1] = out[2] = float(1.0 / 3.0);
        return;
    }
    const float inv = 1 / divisor;
    const f3 s = sub3(p, v[0]);
    float b1 = dot3(s, s1) * inv;
    float b2 = dot3(w, cross3(s, e1)) * inv;
    b1 = clampf(b1, 0, 1);
    b2 = clampf(b2, 0, 1);
    if (b1 + b2 > 1) {
        b1 /= b1 + b2;
        b2 /= b1 + b2;
    }
    out[0] = 1.0f - b1 - b2;
    out[1] = b1;
    out[2] = b2;
}
// util/sampling.cu:41-103
PBRTM_HD void invert_spherical_triangle(const f3 v[3], f3 p, f3 w, float *u0, float *u1) {
    const f3 a = norm3(sub3(v[0], p)), b = norm3(sub3(v[1], p)), c = norm3(sub3(v[2], p));
    f3 n_ab = cross3(a, b), n_bc = cross3(b, c), n_ca = cross3(c, a);
    if (sqlen3(n_ab) == 0 || sqlen3(n_bc) == 0 || sqlen3(n_ca) == 0) {
        *u0 = *u1 = 0;
        return;
    }
    n_ab = norm3(n_ab);
    n_bc = norm3(n_bc);
    n_ca = norm3(n_ca);
    const float alpha = angle_between(n_ab, neg3(n_ca));
    const float beta = angle_between(n_bc, neg3(n_ab));
    const float gamma = angle_between(n_ca, neg3(n_bc));
    f3 cp = norm3(cross3(cross3(b, w), cross3(c, a)));
    if (dot3(cp, add3(a, c)) < 0) cp = neg3(cp);
    float uu0;
    if (dot3(a, cp) > 0.99999847691f) {
        uu0 = 0;
    } else {
        f3 n_cpb = cross3(cp, b), n_acp = cross3(a, cp);
        if (sqlen3(n_cpb) == 0 || sqlen3(n_acp) == 0) {
            *u0 = *u1 = 0.5f;
            return;
        }
        n_cpb = norm3(n_cpb);
        n_acp = norm3(n_acp);
        const float Ap = alpha + angle_between(n_ab, n_cpb) + angle_between(n_acp, neg3(n_cpb)) - kPi;
        const float A = alpha + beta + gamma - kPi;
        uu0 = Ap / A;
    }
    const float uu1 = (1 - dot3(w, b)) / (1 - dot3(cp, b));
    *u0 = clampf(uu0, 0, 1);
    *u1 = clampf(uu1, 0, 1);
}

// Interval point (Point3fi) helpers.
struct P3i {
    Ival x, y, z;
};
PBRTM_HD f3 p3i_mid(const P3i &p) { return {ival_mid(p.x), ival_mid(p.y), ival_mid(p.z)}; }
PBRTM_HD f3 p3i_err(const P3i &p) { return {ival_half(p.x), ival_half(p.y), ival_half(p.z)}; }
PBRTM_HD P3i p3i_make(f3 p, f3 e) { return {ival(p.x, e.x), ival(p.y, e.y), ival(p.z, e.z)}; }
PBRTM_HD P3i p3i_exact(f3 p) { return {{p.x, p.x}, {p.y, p.y}, {p.z, p.z}}; }
PBRTM_HD float round_away(float po, float off) {
    if (off > 0) return next_up(po);
    if (off < 0) return next_down(po);
    return po;
}
// Ray::offset_ray_origin (base/ray.h:22-43): |n . err|.
PBRTM_HD f3 offset_ray_origin(const P3i &pi, f3 n, f3 w) {
    const float d = absdot3(n, p3i_err(pi));
    f3 off = mul3(d, n);
    if (dot3(n, w) < 0) off = neg3(off);
    const f3 po = add3(p3i_mid(pi), off);
    return {round_away(po.x, off.x), round_away(po.y, off.y), round_away(po.z, off.z)};
}
// ShapeSampleContext::offset_ray_origin (base/shape.h:31-53): |n| . err.
PBRTM_HD f3 ctx_offset_origin(const P3i &pi, f3 n, f3 w) {
    const float d = dot3(abs3(n), p3i_err(pi));
    f3 off = mul3(d, n);
    if (dot3(n, w) < 0) off = neg3(off);
    const f3 po = add3(p3i_mid(pi), off);
    return {round_away(po.x, off.x), round_away(po.y, off.y), round_away(po.z, off.z)};
}

// ------------------------------------------------------------ BSDFs in the local frame
// util/scattering.h:82-102
PBRTM_HD float fr_dielectric(float cos_i, float eta) {
    cos_i = clampf(cos_i, -1, 1);
    if (cos_i < 0) {
        eta = 1 / eta;
        cos_i = -cos_i;
    }
    const float sin2_i = 1 - sqrf(cos_i);
    const float sin2_t = sin2_i / sqrf(eta);
    if (sin2_t >= 1) return 1.0f;
    const float cos_t = safe_sqrtf(1 - sin2_t);
    const float r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    const float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (sqrf(r_parl) + sqrf(r_perp)) / 2;
}
// util/scattering.h:15-44
PBRTM_HD bool refract(f3 wi, f3 n, float eta, float *etap, f3 *wt) {
    float cos_i = dot3(n, wi);
    if (cos_i < 0) {
        eta = 1 / eta;
        cos_i = -cos_i;
        n = neg3(n);
    }
    const float sin2_i = smax(0.0f, 1 - sqrf(cos_i));
    const float sin2_t = sin2_i / sqrf(eta);
    if (sin2_t >= 1) return false;
    const float cos_t = sqrtf(1 - sin2_t);
    const f3 a = neg3(wi);
    *wt = add3(f3{a.x / eta, a.y / eta, a.z / eta}, mul3(cos_i / eta - cos_t, n));
    *etap = eta;
    return true;
}
PBRTM_HD f3 reflect3(f3 wo, f3 n) { return add3(neg3(wo), mul3(2 * dot3(wo, n), n)); }
PBRTM_HD f3 face_fwd(f3 v, f3 w) { return dot3(v, w) < 0.0f ? neg3(v) : v; }
PBRTM_HD bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0; }
PBRTM_HD float cos2_th(f3 w) { return sqrf(w.z); }
PBRTM_HD float sin2_th(f3 w) { return smax(0.0f, float(1.0 - double(cos2_th(w)))); }
PBRTM_HD float tan2_th(f3 w) { return sin2_th(w) / cos2_th(w); }
PBRTM_HD float cos_ph(f3 w) {
    const float s = sqrtf(sin2_th(w));
    return (s == 0) ? 1 : clampf(w.x / s, -1, 1);
}
PBRTM_HD float sin_ph(f3 w) {
    const float s = sqrtf(sin2_th(w));
    return (s == 0) ? 0 : clampf(w.y / s, -1, 1);
}

// Trowbridge-Reitz distribution (util/scattering.h:105-229).
struct TR {
    float ax, ay;
    PBRTM_HD bool smooth() const { return smax(ax, ay) < 1e-3f; }
    PBRTM_HD float D(f3 wm) const {
        const float t2 = tan2_th(wm);
        if (isinf(t2)) return 0;
        const float cos4 = sqrf(cos2_th(wm));
        if (cos4 < 1e-16f) return 0;
        const float e = t2 * (sqrf(cos_ph(wm) / ax) + sqrf(sin_ph(wm) / ay));
        return 1 / (kPi * ax * ay * cos4 * sqrf(1 + e));
    }
    PBRTM_HD float lambda(f3 w) const {
        const float t2 = tan2_th(w);
        if (isinf(t2)) return 0;
        const float a2 = sqrf(cos_ph(w) * ax) + sqrf(sin_ph(w) * ay);
        return (sqrtf(1 + a2 * t2) - 1) / 2;
    }
    PBRTM_HD float G1(f3 w) const { return 1 / (1 + lambda(w)); }
    PBRTM_HD float G(f3 wo, f3 wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    PBRTM_HD float pdf(f3 w, f3 wm) const { return G1(w) / fabsf(w.z) * D(wm) * absdot3(w, wm); }
    PBRTM_HD f3 sample_wm(f3 w, float u0, float u1) const {
        f3 wh = norm3(f3{ax * w.x, ay * w.y, w.z});
        if (wh.z < 0) wh = neg3(wh);
        const f3 T1 = (wh.z < 0.99999f) ? norm3(cross3(f3{0.f, 0.f, 1.f}, wh)) : f3{1.f, 0.f, 0.f};
        const f3 T2 = cross3(wh, T1);
        const float r = sqrtf(u0), th = 2 * kPi * u1;  // sampling.h:85-90
        const float px = r * tcos(th);
        float py = r * tsin(th);
        const float h = sqrtf(1 - sqrf(px));
        py = lerpf((1 + wh.z) / 2, h, py);
        const float pz = sqrtf(smax(0.0f, 1 - (px * px + py * py)));
        const f3 nh = add3(add3(mul3(px, T1), mul3(py, T2)), mul3(pz, wh));
        return norm3(f3{ax * nh.x, ay * nh.y, smax(1e-6f, nh.z)});
    }
};

enum : uint32_t { kBxDiffuse = 0, kBxConductor = 1, kBxDielectric = 2 };
// BxDFFlags (base/bxdf.h): reflection 1, transmission 2, diffuse 4, glossy 8, specular 16
struct Bxdf {
    uint32_t type;
    S4 r, eta, k;  // diffuse reflectance; conductor eta, k
    float eta_d;   // dielectric
    TR tr;
    PBRTM_HD uint32_t flags() const {
        if (type == kBxDiffuse) return s4pos(r) ? (1u | 4u) : 0u;
        if (type == kBxConductor) return tr.smooth() ? (1u | 16u) : (1u | 8u);
        return (eta_d == 1 ? 2u : (1u | 2u)) | (tr.smooth() ? 16u : 8u);
    }
    // f; zero where the reference returns an unset spectrum (callers test is_positive)
    PBRTM_HD S4 f(f3 wo, f3 wi) const {
        if (type == kBxDiffuse) return same_hemi(wo, wi) ? s4scale(r, kInvPi) : s4(0.f);
        if (type == kBxConductor) {  // bxdfs/conductor_bxdf.cu:54-83
            if (!same_hemi(wo, wi) || tr.smooth()) return s4(0.f);
            const float co = fabsf(wo.z), ci = fabsf(wi.z);
            if (ci == 0 || co == 0) return s4(0.f);
            f3 wm = add3(wi, wo);
            if (sqlen3(wm) == 0) return s4(0.f);
            wm = norm3(wm);
            S4 F;
            for (int i = 0; i < kNS; ++i) F.v[i] = fr_complex(absdot3(wo, wm), eta.v[i], k.v[i]);
            return s4div(s4scale(s4scale(F, tr.D(wm)), tr.G(wo, wi)), 4 * ci * co);
        }
        // bxdfs/dielectric_bxdf.cu:113-164
        if (eta_d == 1 || tr.smooth()) return s4(0.f);
        const float co = wo.z, ci = wi.z;
        const bool refl = ci * co > 0;
        float etap = 1;
        if (!refl) etap = co > 0 ? eta_d : (1 / eta_d);
        f3 wm = add3(mul3(etap, wi), wo);
        if (ci == 0 || co == 0 || sqlen3(wm) == 0) return s4(0.f);
        wm = norm3(face_fwd(wm, f3{0.f, 0.f, 1.f}));
        if (dot3(wm, wi) * ci < 0 || dot3(wm, wo) * co < 0) return s4(0.f);
        const float F = fr_dielectric(dot3(wo, wm), eta_d);
        if (refl) return s4(tr.D(wm) * tr.G(wo, wi) * F / fabsf(4 * ci * co));
        const float denom = sqrf(dot3(wi, wm) + dot3(wo, wm) / etap) * ci * co;
        float ft = tr.D(wm) * (1 - F) * tr.G(wo, wi) * fabsf(dot3(wi, wm) * dot3(wo, wm) / denom);
        ft /= sqrf(etap);
        return s4(ft);
    }
    PBRTM_HD float pdf(f3 wo, f3 wi) const {
        if (type == kBxDiffuse) return same_hemi(wo, wi) ? fabsf(wi.z) / kPi : 0.f;
        if (type == kBxConductor) {  // conductor_bxdf.cu:85-106
            if (!same_hemi(wo, wi) || tr.smooth()) return 0;
            f3 wm = add3(wo, wi);
            if (sqlen3(wm) == 0) return 0;
            wm = face_fwd(norm3(wm), f3{0.f, 0.f, 1.f});
            return tr.pdf(wo, wm) / (4 * absdot3(wo, wm));
        }
        // dielectric_bxdf.cu:166-230
        if (eta_d == 1 || tr.smooth()) return 0;
        const float co = wo.z, ci = wi.z;
        const bool refl = ci * co > 0;
        float etap = 1;
        if (!refl) etap = co > 0 ? eta_d : (1 / eta_d);
        f3 wm = add3(mul3(etap, wi), wo);
        if (ci == 0 || co == 0 || sqlen3(wm) == 0) return 0;
        wm = norm3(face_fwd(wm, f3{0.f, 0.f, 1.f}));
        if (dot3(wm, wi) * ci < 0 || dot3(wm, wo) * co < 0) return 0;
        const float R = fr_dielectric(dot3(wo, wm), eta_d), T = 1 - R;
        const float pr = R, pt = T;
        if (pr == 0 && pt == 0) return 0;
        if (refl) return tr.pdf(wo, wm) / (4 * absdot3(wo, wm)) * pr / (pr + pt);
        const float denom = sqrf(dot3(wi, wm) + dot3(wo, wm) / etap);
        const float dwm_dwi = absdot3(wi, wm) / denom;
        return tr.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
    }
};

struct BxSample {
    S4 f;
    f3 wi;
    float pdf, eta;
    bool specular, transmission;
};
// BxDF::sample_f of each lobe plus BSDF::sample_f's checks (base/bsdf.cu:57-80); wo local.
PBRTM_HD bool bxdf_sample(const Bxdf &b, f3 wo, float uc, float u0, float u1, BxSample *out) {
    if (wo.z == 0 || b.flags() == 0) return false;
    BxSample bs;
    bs.eta = 1;
    bs.transmission = false;
    bs.specular = false;
    if (b.type == kBxDiffuse) {
        f3 wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0) wi.z *= -1;
        bs.pdf = fabsf(wi.z) / kPi;
        bs.f = s4scale(b.r, kInvPi);
        bs.wi = wi;
    } else if (b.type == kBxConductor) {
        if (b.tr.smooth()) {
            const f3 wi = {-wo.x, -wo.y, wo.z};
            const float c = fabsf(wi.z);
            S4 F;
            for (int i = 0; i < kNS; ++i) F.v[i] = fr_complex(c, b.eta.v[i], b.k.v[i]);
            bs.f = s4div(F, c);
            bs.wi = wi;
            bs.pdf = 1;
            bs.specular = true;
        } else {  // conductor_bxdf.cu:24-52
            const f3 wm = b.tr.sample_wm(wo, u0, u1);
            const f3 wi = reflect3(wo, wm);
            if (!same_hemi(wo, wi)) return false;
            const float pdf = b.tr.pdf(wo, wm) / (4 * absdot3(wo, wm));
            const float co = fabsf(wo.z), ci = fabsf(wi.z);
            if (ci == 0 || co == 0) return false;
            S4 F;
            for (int i = 0; i < kNS; ++i) F.v[i] = fr_complex(absdot3(wo, wm), b.eta.v[i], b.k.v[i]);
            bs.f = s4div(s4scale(s4scale(F, b.tr.D(wm)), b.tr.G(wo, wi)), 4 * ci * co);
            bs.wi = wi;
            bs.pdf = pdf;
        }
    } else if (b.eta_d == 1 || b.tr.smooth()) {  // dielectric_bxdf.cu:4-49
        const float R = fr_dielectric(wo.z, b.eta_d), T = 1 - R;
        const float pr = R, pt = T;
        if (pr == 0 && pt == 0) return false;
        if (uc < pr / (pr + pt)) {
            const f3 wi = {-wo.x, -wo.y, wo.z};
            bs.f = s4(R / fabsf(wi.z));
            bs.wi = wi;
            bs.pdf = pr / (pr + pt);
            bs.specular = true;
        } else {
            f3 wi;
            float etap;
            if (!refract(wo, f3{0.f, 0.f, 1.f}, b.eta_d, &etap, &wi)) return false;
            bs.f = s4div(s4(T / fabsf(wi.z)), sqrf(etap));
            bs.wi = wi;
            bs.pdf = pt / (pr + pt);
            bs.specular = true;
            bs.transmission = true;
            bs.eta = etap;
        }
    } else {  // rough dielectric (dielectric_bxdf.cu:50-110)
        const f3 wm = b.tr.sample_wm(wo, u0, u1);
        const float R = fr_dielectric(dot3(wo, wm), b.eta_d), T = 1 - R;
        const float pr = R, pt = T;
        if (pr == 0 && pt == 0) return false;
        if (uc < pr / (pr + pt)) {
            const f3 wi = reflect3(wo, wm);
            if (!same_hemi(wo, wi)) return false;
            bs.pdf = b.tr.pdf(wo, wm) / (4 * absdot3(wo, wm)) * pr / (pr + pt);
            bs.f = s4(b.tr.D(wm) * b.tr.G(wo, wi) * R / (4 * wi.z * wo.z));
            bs.wi = wi;
        } else {
            float etap;
            f3 wi;
            const bool tir = !refract(wo, wm, b.eta_d, &etap, &wi);
            if (same_hemi(wo, wi) || wi.z == 0 || tir) return false;
            const float denom = sqrf(dot3(wi, wm) + dot3(wo, wm) / etap);
            const float dwm_dwi = absdot3(wi, wm) / denom;
            bs.pdf = b.tr.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
            const float ft = T * b.tr.D(wm) * b.tr.G(wo, wi) * fabsf(dot3(wi, wm) * dot3(wo, wm) / (wi.z * wo.z * denom));
            bs.f = s4div(s4(ft), sqrf(etap));
            bs.wi = wi;
            bs.transmission = true;
            bs.eta = etap;
        }
    }
    if (!s4pos(bs.f) || bs.pdf == 0 || bs.wi.z == 0) return false;
    *out = bs;
    return true;
}

}  // namespace pt
}  // namespace pbrtm
