// TEST INFRASTRUCTURE ONLY — part of the CPU oracle (see oracle_pt.h).
//
// Host restatement of the reference's analytic shapes and instancing: Interval
// arithmetic (util/interval.h), the Transform operations they use
// (euclidean_space/transform.h, transform.cu), Sphere (shapes/sphere.{h,cu}),
// Disk (shapes/disk.{h,cu}) and TransformedPrimitive
// (primitives/transformed_primitive.cu). Expressions keep the reference's form
// and operation order (written with operators, as the reference writes them).
//
// Directed rounding: the reference's device intrinsics (__fmul_ru and friends,
// util/bit_arithmetic.h:137-287) return the exact result rounded toward +/-inf.
// Here products, quotients and square roots are decided exactly in double
// precision (a product or square of two floats is exact in double; a float
// quotient q is above a / b exactly when q * b, exact in double, exceeds a for
// b > 0), and sums use the TwoSum remainder (oracle_math.h).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

#include "oracle_math.h"
#include "oracle_tri.h"

namespace orc {
namespace quad {

inline float qsq(float v) { return v * v; }  // sqr (util/basic_math.h)
inline float step_down(float f) { return std::nextafter(f, -std::numeric_limits<float>::infinity()); }
inline float step_up(float f) { return std::nextafter(f, std::numeric_limits<float>::infinity()); }

inline float mul_rd(float a, float b) {
    const double e = double(a) * double(b);
    const float f = float(e);
    if (!std::isfinite(f)) return f;
    return double(f) > e ? step_down(f) : f;
}
inline float mul_ru(float a, float b) {
    const double e = double(a) * double(b);
    const float f = float(e);
    if (!std::isfinite(f)) return f;
    return double(f) < e ? step_up(f) : f;
}
// q = RN(a / b) against the exact quotient: sign of (q b - a) / b, with q b exact in double.
inline int quot_side(float q, float a, float b) {
    const double r = double(q) * double(b) - double(a);  // exact: |q b| and |a| share the exponent range
    if (r == 0) return 0;
    return ((r > 0) == (b > 0)) ? 1 : -1;  // +1: q above the exact quotient
}
inline float div_rd(float a, float b) {
    const float q = a / b;
    if (!std::isfinite(q)) return q;
    return quot_side(q, a, b) > 0 ? step_down(q) : q;
}
inline float div_ru(float a, float b) {
    const float q = a / b;
    if (!std::isfinite(q)) return q;
    return quot_side(q, a, b) < 0 ? step_up(q) : q;
}
inline float sqrt_rd(float a) {
    const float s = std::sqrt(a);
    if (!std::isfinite(s)) return s;
    return double(s) * double(s) > double(a) ? step_down(s) : s;
}
inline float sqrt_ru(float a) {
    const float s = std::sqrt(a);
    if (!std::isfinite(s)) return s;
    return double(s) * double(s) < double(a) ? step_up(s) : s;
}

// util/interval.h
struct Ivl {
    float low, high;
    Ivl() : low(NAN), high(NAN) {}
    explicit Ivl(float v) : low(v), high(v) {}
    Ivl(float a, float b) : low(std::min(a, b)), high(std::max(a, b)) {}
    static Ivl from_value_and_error(float v, float err) {
        Ivl i;
        if (err == 0) {
            i.low = i.high = v;
        } else {
            i.low = add_round_down(v, -err);
            i.high = add_round_up(v, err);
        }
        return i;
    }
    float midpoint() const { return (low + high) / 2; }
    float width() const { return high - low; }
    bool cover(float v) const { return v >= low && v <= high; }
    bool operator==(const Ivl &i) const { return low == i.low && high == i.high; }
    Ivl operator+(const Ivl &i) const { return {add_round_down(low, i.low), add_round_up(high, i.high)}; }
    Ivl operator+(float f) const { return *this + Ivl(f); }
    Ivl operator-(const Ivl &i) const { return {add_round_down(low, -i.high), add_round_up(high, -i.low)}; }
    Ivl operator*(float f) const {
        return f > 0.0 ? Ivl(mul_rd(f, low), mul_ru(f, high)) : Ivl(mul_rd(f, high), mul_ru(f, low));
    }
    Ivl operator*(const Ivl &i) const {
        const float lp[4] = {mul_rd(low, i.low), mul_rd(high, i.low), mul_rd(low, i.high), mul_rd(high, i.high)};
        const float hp[4] = {mul_ru(low, i.low), mul_ru(high, i.low), mul_ru(low, i.high), mul_ru(high, i.high)};
        return {std::min({lp[0], lp[1], lp[2], lp[3]}), std::max({hp[0], hp[1], hp[2], hp[3]})};
    }
    Ivl operator/(const Ivl &i) const {
        if (i.cover(0)) return Ivl(-std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity());
        const float lq[4] = {div_rd(low, i.low), div_rd(high, i.low), div_rd(low, i.high), div_rd(high, i.high)};
        const float hq[4] = {div_ru(low, i.low), div_ru(high, i.low), div_ru(low, i.high), div_ru(high, i.high)};
        return {std::min({lq[0], lq[1], lq[2], lq[3]}), std::max({hq[0], hq[1], hq[2], hq[3]})};
    }
    Ivl sqrt() const { return {sqrt_rd(low), sqrt_ru(high)}; }
};
inline Ivl operator*(float f, const Ivl &i) { return i * f; }
inline Ivl sqr(const Ivl &i) {
    float al = std::fabs(i.low), ah = std::fabs(i.high);
    if (al > ah) std::swap(al, ah);
    if (i.cover(0)) return Ivl(0, mul_ru(ah, ah));
    return Ivl(mul_rd(al, al), mul_ru(ah, ah));
}

struct IV3 {  // Point3fi / Vector3fi
    Ivl x, y, z;
    V3 mid() const { return {x.midpoint(), y.midpoint(), z.midpoint()}; }
    V3 err() const { return {x.width() / 2, y.width() / 2, z.width() / 2}; }
};
inline IV3 iv3_exact(V3 p) { return {Ivl(p.x), Ivl(p.y), Ivl(p.z)}; }
inline IV3 iv3_of(V3 p, V3 e) {
    return {Ivl::from_value_and_error(p.x, e.x), Ivl::from_value_and_error(p.y, e.y),
            Ivl::from_value_and_error(p.z, e.z)};
}
inline P3fi to_p3fi(const IV3 &p) { return {{p.x.low, p.x.high}, {p.y.low, p.y.high}, {p.z.low, p.z.high}}; }
inline IV3 from_p3fi(const P3fi &p) { return {Ivl(p.x.lo, p.x.hi), Ivl(p.y.lo, p.y.hi), Ivl(p.z.lo, p.z.hi)}; }

// A Transform: m and inv_m, row-major [4][4] (affine shapes only).
struct Tf {
    float m[4][4], inv[4][4];
};

// Point3<Interval> through transform.h:254-265 (w = 1 for affine matrices).
inline IV3 tf_point(const float (&m)[4][4], const IV3 &p) {
    IV3 r;
    r.x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    r.y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    r.z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    return r;
}
inline V3 tf_point(const float (&m)[4][4], V3 p) {
    return {m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3],
            m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
            m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]};
}
inline V3 tf_vector(const float (&m)[4][4], V3 v) {
    return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}
inline V3 tf_normal(const float (&inv)[4][4], V3 n) {  // transform.h:309-318
    return {inv[0][0] * n.x + inv[1][0] * n.y + inv[2][0] * n.z, inv[0][1] * n.x + inv[1][1] * n.y + inv[2][1] * n.z,
            inv[0][2] * n.x + inv[1][2] * n.y + inv[2][2] * n.z};
}
// Vector3fi of an exact vector (transform.h:267-300).
inline IV3 tf_vector_fi(const float (&m)[4][4], V3 v) {
    const float g = gammaf(3);
    V3 e;
    e.x = g * (std::fabs(m[0][0] * v.x) + std::fabs(m[0][1] * v.y) + std::fabs(m[0][2] * v.z));
    e.y = g * (std::fabs(m[1][0] * v.x) + std::fabs(m[1][1] * v.y) + std::fabs(m[1][2] * v.z));
    e.z = g * (std::fabs(m[2][0] * v.x) + std::fabs(m[2][1] * v.y) + std::fabs(m[2][2] * v.z));
    return iv3_of(tf_vector(m, v), e);
}
// Transform::apply_inverse(Ray, &tMax) (transform.cu:36-55).
inline void tf_ray_inverse(const Tf &t, V3 ro, V3 rd, float *t_max, V3 *o, V3 *d) {
    const auto &im = t.inv;
    const float x = ro.x, y = ro.y, z = ro.z;
    const V3 p = {(im[0][0] * x + im[0][1] * y) + (im[0][2] * z + im[0][3]),
                  (im[1][0] * x + im[1][1] * y) + (im[1][2] * z + im[1][3]),
                  (im[2][0] * x + im[2][1] * y) + (im[2][2] * z + im[2][3])};
    const float g = gammaf(3);
    const V3 e = {g * (std::fabs(im[0][0] * x) + std::fabs(im[0][1] * y) + std::fabs(im[0][2] * z)),
                  g * (std::fabs(im[1][0] * x) + std::fabs(im[1][1] * y) + std::fabs(im[1][2] * z)),
                  g * (std::fabs(im[2][0] * x) + std::fabs(im[2][1] * y) + std::fabs(im[2][2] * z))};
    IV3 oi = iv3_of(p, e);
    const V3 dd = tf_vector(im, rd);
    const float ls = sqlen(dd);
    if (ls > 0) {
        const V3 oe = {oi.x.width() / 2, oi.y.width() / 2, oi.z.width() / 2};
        const float dt = std::fabs(dot(dd, oe)) / ls;
        const V3 off = scale(dt, dd);
        oi.x = oi.x + off.x;
        oi.y = oi.y + off.y;
        oi.z = oi.z + off.z;
        *t_max -= dt;
    }
    *o = oi.mid();
    *d = dd;
}

// The analytic primitive as the oracle keeps it.
struct Quadric {
    uint32_t type;  // 1 sphere, 2 disk, 3 instanced triangle
    Tf rfo;         // render_from_object (instances: render_from_instance)
    bool reverse, flip;
    float radius, z_min, z_max, theta_z_min, theta_z_max, phi_max, height, inner_radius;
    V3 p[3];        // instances: the base triangle
};

struct QuadricHit {
    float t;
    V3 p;  // object-space point (sphere, disk) or base barycentrics (instance)
};

inline float tan2_phi(V3 p, float pi) {  // std::atan2 rounded from double, shifted to [0, 2 pi)
    float phi = float(std::atan2(double(p.y), double(p.x)));
    if (phi < 0) phi += 2 * pi;
    return phi;
}

// Sphere::basic_intersect (shapes/sphere.cu:49-150).
inline bool sphere_intersect(const Quadric &s, V3 ro, V3 rd, float tMax, QuadricHit *out, float pi) {
    const IV3 oi = tf_point(s.rfo.inv, iv3_exact(ro));
    const IV3 di = tf_vector_fi(s.rfo.inv, rd);
    const Ivl a = sqr(di.x) + sqr(di.y) + sqr(di.z);
    const Ivl b = 2 * (di.x * oi.x + di.y * oi.y + di.z * oi.z);
    const Ivl c = sqr(oi.x) + sqr(oi.y) + sqr(oi.z) - sqr(Ivl(s.radius));
    const Ivl f = b / (2 * a);
    const IV3 v = {oi.x - di.x * f, oi.y - di.y * f, oi.z - di.z * f};
    const Ivl length = (v.x * v.x + v.y * v.y + v.z * v.z).sqrt();
    const Ivl discrim = 4 * a * (Ivl(s.radius) + length) * (Ivl(s.radius) - length);
    if (discrim.low < 0) return false;
    const Ivl root = discrim.sqrt();
    Ivl q = b.midpoint() < 0 ? -.5f * (b - root) : -.5f * (b + root);
    Ivl t0 = q / a, t1 = c / q;
    if (t0.low > t1.low) std::swap(t0, t1);
    if (t0.high > tMax || t1.low <= 0) return false;
    Ivl tShapeHit = t0;
    if (tShapeHit.low <= 0) {
        tShapeHit = t1;
        if (tShapeHit.high > tMax) return false;
    }
    V3 pHit;
    float phi;
    auto refine = [&]() {
        pHit = add(oi.mid(), scale(tShapeHit.midpoint(), di.mid()));
        pHit = scale(s.radius / std::sqrt(qsq(pHit.x) + qsq(pHit.y) + qsq(pHit.z)), pHit);
        if (pHit.x == 0 && pHit.y == 0) pHit.x = 1e-5f * s.radius;
        phi = tan2_phi(pHit, pi);
    };
    auto clipped = [&]() {
        return (s.z_min > -s.radius && pHit.z < s.z_min) || (s.z_max < s.radius && pHit.z > s.z_max) ||
               phi > s.phi_max;
    };
    refine();
    if (clipped()) {
        if (tShapeHit == t1) return false;
        if (t1.high > tMax) return false;
        tShapeHit = t1;
        refine();
        if (clipped()) return false;
    }
    out->t = tShapeHit.midpoint();
    out->p = pHit;
    return true;
}

// Disk::basic_intersect (shapes/disk.cu:96-138).
inline bool disk_intersect(const Quadric &s, V3 ro, V3 rd, float tMax, QuadricHit *out, float pi) {
    const IV3 oi = tf_point(s.rfo.inv, iv3_exact(ro));
    const IV3 di = tf_vector_fi(s.rfo.inv, rd);
    if (di.z.midpoint() == 0) return false;
    const float tShapeHit = (s.height - oi.z.midpoint()) / di.z.midpoint();
    if (tShapeHit <= 0 || tShapeHit >= tMax) return false;
    const V3 pHit = add(oi.mid(), scale(tShapeHit, di.mid()));
    const float dist2 = qsq(pHit.x) + qsq(pHit.y);
    if (dist2 > qsq(s.radius) || dist2 < qsq(s.inner_radius)) return false;
    if (tan2_phi(pHit, pi) > s.phi_max) return false;
    out->t = tShapeHit;
    out->p = pHit;
    return true;
}

// TransformedPrimitive::intersect (primitives/transformed_primitive.cu:21-34).
inline bool instance_intersect(const Quadric &s, V3 ro, V3 rd, float tMax, QuadricHit *out) {
    V3 o, d;
    float tm = tMax;
    tf_ray_inverse(s.rfo, ro, rd, &tm, &o, &d);
    TriHit h;
    if (!intersect_triangle(o, d, tm, s.p[0], s.p[1], s.p[2], &h)) return false;
    out->t = h.t;
    out->p = {h.b0, h.b1, h.b2};
    return true;
}

inline bool quadric_intersect(const Quadric &s, V3 ro, V3 rd, float tMax, QuadricHit *out, float pi) {
    if (s.type == 1) return sphere_intersect(s, ro, rd, tMax, out, pi);
    if (s.type == 2) return disk_intersect(s, ro, rd, tMax, out, pi);
    return instance_intersect(s, ro, rd, tMax, out);
}

// The members of a SurfaceInteraction the renderer reads.
struct QSurf {
    P3fi pi;
    V3 n, ns, dpdu;
};

// SurfaceInteraction(pi, uv, wo, dpdu, dpdv, dndu, dndv, flip) (base/interaction.h:97-115),
// then render_from_object(si) (transform.cu:57-97).
inline QSurf to_render(const Quadric &s, const IV3 &pi_obj, V3 dpdu, V3 dpdv) {
    V3 n = normalize(cross(dpdu, dpdv));
    V3 ns = n;
    if (s.flip) {
        n = scale(-1.f, n);
        ns = scale(-1.f, ns);
    }
    QSurf r;
    r.pi = to_p3fi(tf_point(s.rfo.m, pi_obj));
    r.n = normalize(tf_normal(s.rfo.inv, n));
    r.dpdu = tf_vector(s.rfo.m, dpdu);
    r.ns = normalize(tf_normal(s.rfo.inv, ns));
    r.n = dot(r.n, r.ns) < 0.0 ? neg(r.n) : r.n;
    r.ns = dot(r.ns, r.n) < 0.0 ? neg(r.ns) : r.ns;
    return r;
}

// Sphere::interaction_from_intersection (shapes/sphere.h:75-135).
inline QSurf sphere_surface(const Quadric &s, V3 pHit) {
    const float cosTheta = pHit.z / s.radius;
    const float zRadius = std::sqrt(qsq(pHit.x) + qsq(pHit.y));
    const float cosPhi = pHit.x / zRadius, sinPhi = pHit.y / zRadius;
    const V3 dpdu = {-s.phi_max * pHit.y, s.phi_max * pHit.x, 0};
    const float sinTheta = std::sqrt(std::max(0.f, 1 - qsq(cosTheta)));
    const V3 dpdv = scale(s.theta_z_max - s.theta_z_min, V3{pHit.z * cosPhi, pHit.z * sinPhi, -s.radius * sinTheta});
    const V3 pError = scale(gammaf(5), vabs(pHit));
    return to_render(s, iv3_of(pHit, pError), dpdu, dpdv);
}

// Disk::interaction_from_intersection (shapes/disk.cu:140-170).
inline QSurf disk_surface(const Quadric &s, V3 pHit) {
    const float rHit = std::sqrt(qsq(pHit.x) + qsq(pHit.y));
    const V3 dpdu = {-s.phi_max * pHit.y, s.phi_max * pHit.x, 0};
    const V3 v = {pHit.x, pHit.y, 0};
    const float k = s.inner_radius - s.radius;
    const V3 dpdv = {(v.x * k) / rHit, (v.y * k) / rHit, (v.z * k) / rHit};
    pHit.z = s.height;
    return to_render(s, iv3_exact(pHit), dpdu, dpdv);
}

// Triangle::interaction_from_intersection (shapes/triangle.cu:330-470) of the
// instance's triangle, then render_from_instance.
inline QSurf instance_surface(const Quadric &s, V3 bary) {
    const V3 p0 = s.p[0], p1 = s.p[1], p2 = s.p[2];
    const V3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
    const float du02[2] = {-1.f, -1.f}, du12[2] = {0.f, -1.f};
    const float det = dop(du02[0], du12[1], du02[1], du12[0]);
    V3 dpdu, dpdv;
    const bool degenerate = std::fabs(det) < 1e-9f;
    if (!degenerate) {
        const float invdet = 1 / det;
        dpdu = scale(invdet, V3{dop(du12[1], dp02.x, du02[1], dp12.x), dop(du12[1], dp02.y, du02[1], dp12.y),
                                dop(du12[1], dp02.z, du02[1], dp12.z)});
        dpdv = scale(invdet, V3{dop(du02[0], dp12.x, du12[0], dp02.x), dop(du02[0], dp12.y, du12[0], dp02.y),
                                dop(du02[0], dp12.z, du12[0], dp02.z)});
    }
    if (degenerate || sqlen(cross(dpdu, dpdv)) == 0) coordinate_system(normalize(cross(sub(p2, p0), sub(p1, p0))), &dpdu, &dpdv);
    const HitGeom g = hit_geometry(TriHit{bary.x, bary.y, bary.z, 0}, p0, p1, p2);
    V3 n = normalize(cross(dp02, dp12));
    if (s.flip) n = neg(n);
    QSurf r;
    r.pi = to_p3fi(tf_point(s.rfo.m, from_p3fi(g.pi)));
    r.n = normalize(tf_normal(s.rfo.inv, n));
    r.dpdu = tf_vector(s.rfo.m, dpdu);
    r.ns = normalize(tf_normal(s.rfo.inv, n));
    r.n = dot(r.n, r.ns) < 0.0 ? neg(r.n) : r.n;
    r.ns = dot(r.ns, r.n) < 0.0 ? neg(r.ns) : r.ns;
    return r;
}

inline QSurf quadric_surface(const Quadric &s, V3 hp) {
    if (s.type == 1) return sphere_surface(s, hp);
    if (s.type == 2) return disk_surface(s, hp);
    return instance_surface(s, hp);
}

inline float sphere_area(const Quadric &s) { return s.phi_max * s.radius * (s.z_max - s.z_min); }
inline float disk_area(const Quadric &s) {
    return float(0.5 * s.phi_max * (qsq(s.radius) - qsq(s.inner_radius)));
}
inline float quadric_area(const Quadric &s) { return s.type == 1 ? sphere_area(s) : disk_area(s); }

}  // namespace quad
}  // namespace orc
