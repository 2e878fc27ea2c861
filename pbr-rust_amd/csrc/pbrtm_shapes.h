// Analytic shapes (spheres and disks) and instanced triangles
// (TransformedPrimitive): interval arithmetic with directed rounding, the
// reference's Transform operations, ray intersection, the surface interaction
// in render space and the area-light sampling of spheres and disks.
//
// FP32 operation order follows the reference (cited per function) so that,
// compiled with -ffp-contract=off, results are bit-identical to the host
// restatement. Directed-rounding products, quotients and square roots are the
// exact results rounded toward +inf / -inf (as the device intrinsics the
// reference calls, bit_arithmetic.h:137-287): the correctly rounded result,
// stepped one ulp when the FMA residual shows the exact value lies beyond it.
#pragma once

#include "pt_device.h"

namespace pbrtm {
namespace pt {

constexpr uint32_t kShapeSphere = 1, kShapeDisk = 2, kShapeInstance = 3;

// One analytic primitive, 224 bytes. m / inv: render_from_object and its
// inverse (row-major; affine, last row 0 0 0 1). Instance: m / inv are
// render_from_instance, p0..p2 the base triangle in the definition's render space.
struct DShape {
    uint32_t type;
    uint32_t flip;       // reverse_orientation ^ transform_swaps_handedness (normal flip)
    uint32_t reverse;    // reverse_orientation (sampled light normals)
    uint32_t pad;
    float radius, zmin, zmax, theta_zmin, theta_zmax, phimax, height, inner;
    float m[16], inv[16];
    float p[9];
    float pad2[3];
};
static_assert(sizeof(DShape) == 224, "DShape layout");

// ------------------------------------------------ directed rounding (exact)
PBRTM_HD float mul_up(float a, float b) {
    const float p = a * b;
    if (!isfinite(p)) return p;
    return fmaf(a, b, -p) > 0.f ? next_up(p) : p;
}
PBRTM_HD float mul_down(float a, float b) {
    const float p = a * b;
    if (!isfinite(p)) return p;
    return fmaf(a, b, -p) < 0.f ? next_down(p) : p;
}
PBRTM_HD float div_up(float a, float b) {
    const float q = a / b;
    if (!isfinite(q)) return q;
    const float r = fmaf(-q, b, a);  // a - q b exactly: the quotient's remainder
    return (r != 0.f && ((r > 0.f) == (b > 0.f))) ? next_up(q) : q;
}
PBRTM_HD float div_down(float a, float b) {
    const float q = a / b;
    if (!isfinite(q)) return q;
    const float r = fmaf(-q, b, a);
    return (r != 0.f && ((r > 0.f) != (b > 0.f))) ? next_down(q) : q;
}
PBRTM_HD float sqrt_up(float a) {
    const float s = sqrtf(a);
    if (!isfinite(s)) return s;
    return fmaf(-s, s, a) > 0.f ? next_up(s) : s;
}
PBRTM_HD float sqrt_down(float a) {
    const float s = sqrtf(a);
    if (!isfinite(s)) return s;
    return fmaf(-s, s, a) < 0.f ? next_down(s) : s;
}

// ------------------------------------------------ Interval (util/interval.h)
PBRTM_HD Ival iv_mk(float a, float b) { return {smin(a, b), smax(a, b)}; }  // Interval(low, high)
PBRTM_HD Ival iv_of(float v) { return {v, v}; }
PBRTM_HD Ival iv_add(Ival a, Ival b) { return iv_mk(add_down(a.lo, b.lo), add_up(a.hi, b.hi)); }
PBRTM_HD Ival iv_sub(Ival a, Ival b) { return iv_mk(add_down(a.lo, -b.hi), add_up(a.hi, -b.lo)); }
PBRTM_HD Ival iv_mulf(Ival a, float f) {
    return f > 0.0f ? iv_mk(mul_down(f, a.lo), mul_up(f, a.hi)) : iv_mk(mul_down(f, a.hi), mul_up(f, a.lo));
}
// std::min / std::max over an initializer list: the first of equal values wins.
PBRTM_HD float min4(const float v[4]) {
    float m = v[0];
    for (int i = 1; i < 4; ++i)
        if (v[i] < m) m = v[i];
    return m;
}
PBRTM_HD float max4(const float v[4]) {
    float m = v[0];
    for (int i = 1; i < 4; ++i)
        if (m < v[i]) m = v[i];
    return m;
}
PBRTM_HD Ival iv_mul(Ival a, Ival b) {
    const float lp[4] = {mul_down(a.lo, b.lo), mul_down(a.hi, b.lo), mul_down(a.lo, b.hi), mul_down(a.hi, b.hi)};
    const float hp[4] = {mul_up(a.lo, b.lo), mul_up(a.hi, b.lo), mul_up(a.lo, b.hi), mul_up(a.hi, b.hi)};
    return iv_mk(min4(lp), max4(hp));
}
PBRTM_HD bool iv_covers(Ival a, float v) { return v >= a.lo && v <= a.hi; }
PBRTM_HD Ival iv_div(Ival a, Ival b) {
    if (iv_covers(b, 0.f)) return iv_mk(-__builtin_inff(), __builtin_inff());
    const float lq[4] = {div_down(a.lo, b.lo), div_down(a.hi, b.lo), div_down(a.lo, b.hi), div_down(a.hi, b.hi)};
    const float hq[4] = {div_up(a.lo, b.lo), div_up(a.hi, b.lo), div_up(a.lo, b.hi), div_up(a.hi, b.hi)};
    return iv_mk(min4(lq), max4(hq));
}
PBRTM_HD Ival iv_sqrt(Ival a) { return iv_mk(sqrt_down(a.lo), sqrt_up(a.hi)); }
PBRTM_HD Ival iv_sqr(Ival a) {
    float al = fabsf(a.lo), ah = fabsf(a.hi);
    if (al > ah) {
        const float t = al;
        al = ah;
        ah = t;
    }
    if (iv_covers(a, 0.f)) return iv_mk(0.f, mul_up(ah, ah));
    return iv_mk(mul_down(al, al), mul_up(ah, ah));
}
PBRTM_HD bool iv_eq(Ival a, Ival b) { return a.lo == b.lo && a.hi == b.hi; }
PBRTM_HD float iv_mid(Ival a) { return (a.lo + a.hi) / 2; }

// ------------------------------------------------ Transform (euclidean_space/transform.h)
// Point3<Interval> through the point template (transform.h:254-265), affine.
PBRTM_HD P3i xf_point_iv(const float *m, const P3i &p) {
    P3i r;
    Ival *o[3] = {&r.x, &r.y, &r.z};
    for (int i = 0; i < 3; ++i) {
        Ival s = iv_add(iv_mulf(p.x, m[4 * i]), iv_mulf(p.y, m[4 * i + 1]));
        s = iv_add(s, iv_mulf(p.z, m[4 * i + 2]));
        *o[i] = iv_add(s, iv_of(m[4 * i + 3]));
    }
    return r;
}
// Point3f through the point template (float), affine.
PBRTM_HD f3 xf_point(const float *m, f3 p) {
    float o[3];
    for (int i = 0; i < 3; ++i) o[i] = ((m[4 * i] * p.x + m[4 * i + 1] * p.y) + m[4 * i + 2] * p.z) + m[4 * i + 3];
    return {o[0], o[1], o[2]};
}
// Vector3f (transform.h:302-307).
PBRTM_HD f3 xf_vec(const float *m, f3 v) {
    float o[3];
    for (int i = 0; i < 3; ++i) o[i] = (m[4 * i] * v.x + m[4 * i + 1] * v.y) + m[4 * i + 2] * v.z;
    return {o[0], o[1], o[2]};
}
// Normal3f through the inverse's transpose (transform.h:309-318).
PBRTM_HD f3 xf_normal(const float *inv, f3 n) {
    float o[3];
    for (int i = 0; i < 3; ++i) o[i] = (inv[i] * n.x + inv[4 + i] * n.y) + inv[8 + i] * n.z;
    return {o[0], o[1], o[2]};
}
// Vector3fi of an exact vector (transform.h:267-300, is_exact branch).
PBRTM_HD P3i xf_vec_iv_exact(const float *m, f3 v) {
    const float g3 = gamma_n(3);
    float o[3], e[3];
    for (int i = 0; i < 3; ++i) {
        e[i] = g3 * ((fabsf(m[4 * i] * v.x) + fabsf(m[4 * i + 1] * v.y)) + fabsf(m[4 * i + 2] * v.z));
        o[i] = (m[4 * i] * v.x + m[4 * i + 1] * v.y) + m[4 * i + 2] * v.z;
    }
    return p3i_make(f3{o[0], o[1], o[2]}, f3{e[0], e[1], e[2]});
}
// Transform::apply_inverse(Ray, &t_max) (transform.cu:36-55) with apply_inverse of an
// exact Point3fi (transform.h:377-422) and of a Vector3f (:354-363).
PBRTM_HD void xf_ray_inverse(const float *inv, f3 o, f3 d, float *t_max, f3 *o_out, f3 *d_out) {
    const float g3 = gamma_n(3);
    float po[3], pe[3];
    for (int i = 0; i < 3; ++i) {
        const float *r = inv + 4 * i;
        po[i] = (r[0] * o.x + r[1] * o.y) + (r[2] * o.z + r[3]);
        pe[i] = g3 * ((fabsf(r[0] * o.x) + fabsf(r[1] * o.y)) + fabsf(r[2] * o.z));
    }
    P3i oi = p3i_make(f3{po[0], po[1], po[2]}, f3{pe[0], pe[1], pe[2]});
    const f3 dd = xf_vec(inv, d);
    const float ls = sqlen3(dd);
    if (ls > 0) {
        const f3 oe = {(oi.x.hi - oi.x.lo) / 2, (oi.y.hi - oi.y.lo) / 2, (oi.z.hi - oi.z.lo) / 2};
        const float dt = absdot3(dd, oe) / ls;
        oi.x = iv_add(oi.x, iv_of(dd.x * dt));
        oi.y = iv_add(oi.y, iv_of(dd.y * dt));
        oi.z = iv_add(oi.z, iv_of(dd.z * dt));
        *t_max -= dt;
    }
    *o_out = p3i_mid(oi);
    *d_out = dd;
}

// ------------------------------------------------ intersection
struct QHit {
    float t;
    f3 p;  // object-space hit point (quadrics), or the base triangle's barycentrics (instances)
};

PBRTM_HD float phi_of(f3 p) {  // atan2 in [0, 2 pi)
    float phi = tatan2(p.y, p.x);
    if (phi < 0) phi += 2 * kPi;
    return phi;
}

// Sphere::basic_intersect (shapes/sphere.cu:49-150).
PBRTM_HD bool sphere_hit(const DShape &s, f3 ro, f3 rd, float t_max, QHit *out) {
    const P3i oi = xf_point_iv(s.inv, p3i_exact(ro));
    const P3i di = xf_vec_iv_exact(s.inv, rd);
    const Ival a = iv_add(iv_add(iv_sqr(di.x), iv_sqr(di.y)), iv_sqr(di.z));
    const Ival b = iv_mulf(iv_add(iv_add(iv_mul(di.x, oi.x), iv_mul(di.y, oi.y)), iv_mul(di.z, oi.z)), 2);
    const Ival c = iv_sub(iv_add(iv_add(iv_sqr(oi.x), iv_sqr(oi.y)), iv_sqr(oi.z)), iv_sqr(iv_of(s.radius)));
    const Ival f = iv_div(b, iv_mulf(a, 2));
    const Ival vx = iv_sub(oi.x, iv_mul(di.x, f)), vy = iv_sub(oi.y, iv_mul(di.y, f)),
               vz = iv_sub(oi.z, iv_mul(di.z, f));
    const Ival len = iv_sqrt(iv_add(iv_add(iv_mul(vx, vx), iv_mul(vy, vy)), iv_mul(vz, vz)));
    const Ival discrim =
        iv_mul(iv_mul(iv_mulf(a, 4), iv_add(iv_of(s.radius), len)), iv_sub(iv_of(s.radius), len));
    if (discrim.lo < 0) return false;
    const Ival root = iv_sqrt(discrim);
    const Ival q = iv_mid(b) < 0 ? iv_mulf(iv_sub(b, root), -.5f) : iv_mulf(iv_add(b, root), -.5f);
    Ival t0 = iv_div(q, a), t1 = iv_div(c, q);
    if (t0.lo > t1.lo) {
        const Ival t = t0;
        t0 = t1;
        t1 = t;
    }
    if (t0.hi > t_max || t1.lo <= 0) return false;
    Ival th = t0;
    if (th.lo <= 0) {
        th = t1;
        if (th.hi > t_max) return false;
    }
    const f3 om = p3i_mid(oi), dm = p3i_mid(di);
    auto point_at = [&](Ival t, f3 *p, float *phi) {
        const float tm = iv_mid(t);
        f3 q3 = add3(om, mul3(tm, dm));
        const float dist = sqrtf((q3.x * q3.x + q3.y * q3.y) + q3.z * q3.z);
        q3 = mul3(s.radius / dist, q3);
        if (q3.x == 0 && q3.y == 0) q3.x = 1e-5f * s.radius;
        *p = q3;
        *phi = phi_of(q3);
    };
    auto clipped = [&](f3 p, float phi) {
        return (s.zmin > -s.radius && p.z < s.zmin) || (s.zmax < s.radius && p.z > s.zmax) || phi > s.phimax;
    };
    f3 p;
    float phi;
    point_at(th, &p, &phi);
    if (clipped(p, phi)) {
        if (iv_eq(th, t1)) return false;
        if (t1.hi > t_max) return false;
        th = t1;
        point_at(th, &p, &phi);
        if (clipped(p, phi)) return false;
    }
    out->t = iv_mid(th);
    out->p = p;
    return true;
}

// Disk::basic_intersect (shapes/disk.cu:96-138).
PBRTM_HD bool disk_hit(const DShape &s, f3 ro, f3 rd, float t_max, QHit *out) {
    const P3i oi = xf_point_iv(s.inv, p3i_exact(ro));
    const P3i di = xf_vec_iv_exact(s.inv, rd);
    const float dz = iv_mid(di.z);
    if (dz == 0) return false;
    const float th = (s.height - iv_mid(oi.z)) / dz;
    if (th <= 0 || th >= t_max) return false;
    const f3 p = add3(p3i_mid(oi), mul3(th, p3i_mid(di)));
    const float d2 = p.x * p.x + p.y * p.y;
    if (d2 > s.radius * s.radius || d2 < s.inner * s.inner) return false;
    if (phi_of(p) > s.phimax) return false;
    out->t = th;
    out->p = p;
    return true;
}

// TransformedPrimitive::intersect (primitives/transformed_primitive.cu:21-34): the ray
// in the instance's space, the base triangle's test against t_max less the origin shift.
PBRTM_HD bool instance_hit(const DShape &s, f3 ro, f3 rd, float t_max, QHit *out) {
    f3 o, d;
    float tm = t_max;
    xf_ray_inverse(s.inv, ro, rd, &tm, &o, &d);
    TriHit h;
    if (!tri_hit(o, d, tm, f3{s.p[0], s.p[1], s.p[2]}, f3{s.p[3], s.p[4], s.p[5]}, f3{s.p[6], s.p[7], s.p[8]}, &h))
        return false;
    out->t = h.t;
    out->p = f3{h.b0, h.b1, h.b2};
    return true;
}

static __device__ __noinline__ bool shape_hit(const DShape *__restrict__ shapes, uint32_t idx, f3 ro, f3 rd, float t_max,
                                       QHit *out) {
    const DShape &s = shapes[idx];
    if (s.type == kShapeSphere) return sphere_hit(s, ro, rd, t_max, out);
    if (s.type == kShapeDisk) return disk_hit(s, ro, rd, t_max, out);
    return instance_hit(s, ro, rd, t_max, out);
}

// ------------------------------------------------ surface interaction
struct Surf {
    P3i pi;
    f3 n, ns, dpdu;
};

// Transform::operator()(SurfaceInteraction) (transform.cu:57-97) for the members the
// renderer reads: the interval point, n, shading n and shading dpdu, then the
// face-forwarding of n to the shading normal and back.
PBRTM_HD Surf surf_to_render(const DShape &s, const P3i &pi, f3 n, f3 ns, f3 dpdu) {
    Surf r;
    r.pi = xf_point_iv(s.m, pi);
    r.n = norm3(xf_normal(s.inv, n));
    r.dpdu = xf_vec(s.m, dpdu);
    r.ns = norm3(xf_normal(s.inv, ns));
    r.n = face_fwd(r.n, r.ns);
    r.ns = face_fwd(r.ns, r.n);
    return r;
}

// Sphere::interaction_from_intersection (shapes/sphere.h:75-135), in render space.
PBRTM_HD Surf sphere_surf(const DShape &s, f3 p) {
    const float zr = sqrtf(p.x * p.x + p.y * p.y);
    const float cos_phi = p.x / zr, sin_phi = p.y / zr;
    const f3 dpdu = {-s.phimax * p.y, s.phimax * p.x, 0};
    const float cos_t = p.z / s.radius;
    const float sin_t = safe_sqrtf(1 - cos_t * cos_t);
    const f3 dpdv = mul3(s.theta_zmax - s.theta_zmin, f3{p.z * cos_phi, p.z * sin_phi, -s.radius * sin_t});
    f3 n = norm3(cross3(dpdu, dpdv));
    if (s.flip) n = mul3(-1.f, n);
    const f3 perr = mul3(gamma_n(5), abs3(p));
    return surf_to_render(s, p3i_make(p, perr), n, n, dpdu);
}

// Disk::interaction_from_intersection (shapes/disk.cu:140-170), in render space.
PBRTM_HD Surf disk_surf(const DShape &s, f3 p) {
    const float rh = sqrtf(p.x * p.x + p.y * p.y);
    const f3 dpdu = {-s.phimax * p.y, s.phimax * p.x, 0};
    const float k = s.inner - s.radius;
    const f3 dpdv = {(p.x * k) / rh, (p.y * k) / rh, (0.f * k) / rh};
    f3 n = norm3(cross3(dpdu, dpdv));
    if (s.flip) n = mul3(-1.f, n);
    return surf_to_render(s, p3i_exact(f3{p.x, p.y, s.height}), n, n, dpdu);
}

// Triangle::interaction_from_intersection (shapes/triangle.cu:330-470; default uv, no
// mesh normals) of the base triangle, then render_from_instance.
PBRTM_HD Surf instance_surf(const DShape &s, f3 b) {
    const f3 p0 = {s.p[0], s.p[1], s.p[2]}, p1 = {s.p[3], s.p[4], s.p[5]}, p2 = {s.p[6], s.p[7], s.p[8]};
    const f3 dp02 = sub3(p0, p2), dp12 = sub3(p1, p2);
    f3 dpdu = {dop(-1.f, dp02.x, -1.f, dp12.x), dop(-1.f, dp02.y, -1.f, dp12.y), dop(-1.f, dp02.z, -1.f, dp12.z)};
    f3 dpdv = {dop(-1.f, dp12.x, 0.f, dp02.x), dop(-1.f, dp12.y, 0.f, dp02.y), dop(-1.f, dp12.z, 0.f, dp02.z)};
    dpdu = mul3(1.0f, dpdu);
    dpdv = mul3(1.0f, dpdv);
    if (sqlen3(cross3(dpdu, dpdv)) == 0) basis(norm3(cross3(sub3(p2, p0), sub3(p1, p0))), &dpdu, &dpdv);
    const HitGeom g = hit_geom(TriHit{b.x, b.y, b.z, 0.f}, p0, p1, p2);
    const f3 n = s.flip ? neg3(g.n) : g.n;
    return surf_to_render(s, P3i{g.x, g.y, g.z}, n, n, dpdu);
}

PBRTM_HD Surf shape_surf(const DShape &s, f3 hp) {
    if (s.type == kShapeSphere) return sphere_surf(s, hp);
    if (s.type == kShapeDisk) return disk_surf(s, hp);
    return instance_surf(s, hp);
}

// ------------------------------------------------ area lights on spheres and disks
PBRTM_HD float sphere_area(const DShape &s) { return s.phimax * s.radius * (s.zmax - s.zmin); }
PBRTM_HD float disk_area(const DShape &s) {
    return float(0.5 * double(s.phimax) * double(s.radius * s.radius - s.inner * s.inner));
}
PBRTM_HD float sqdist3(f3 a, f3 b) { return sqrf(a.x - b.x) + sqrf(a.y - b.y) + sqrf(a.z - b.z); }

struct QSample {
    P3i pi;
    f3 n;
    float pdf;
};

// sample_uniform_disk_concentric (util/sampling.h:61-83).
PBRTM_HD void concentric_disk(float u0, float u1, float *dx, float *dy) {
    const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    *dx = *dy = 0.f;
    if (ox == 0.0f && oy == 0.0f) return;
    const float pi4 = float(double(kPi) / 4.0);
    float r, theta;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = pi4 * (oy / ox);
    } else {
        r = oy;
        theta = float(double(kPi) / 2.0) - pi4 * (ox / oy);
    }
    *dx = r * tcos(theta);
    *dy = r * tsin(theta);
}

// Sphere::sample(u) (shapes/sphere.cu:152-181): area sampling.
PBRTM_HD void sphere_sample_area(const DShape &s, float u0, float u1, QSample *q) {
    const f3 w = uniform_sphere(u0, u1);
    f3 po = {0.f + s.radius * w.x, 0.f + s.radius * w.y, 0.f + s.radius * w.z};
    po = mul3(s.radius / sqrtf((po.x * po.x + po.y * po.y) + po.z * po.z), po);
    const f3 pe = mul3(gamma_n(5), abs3(po));
    f3 n = norm3(xf_normal(s.inv, po));
    if (s.reverse) n = mul3(-1.f, n);
    q->pi = xf_point_iv(s.m, p3i_make(po, pe));
    q->n = n;
    q->pdf = 1.f / sphere_area(s);
}

// Sphere::sample(ctx, u) (shapes/sphere.cu:183-255). cpi / cn: the reference point
// and its geometric normal (ShapeSampleContext).
PBRTM_HD bool sphere_sample(const DShape &s, const P3i &cpi, f3 cn, float u0, float u1, QSample *q) {
    const f3 pc = xf_point(s.m, f3{0.f, 0.f, 0.f});
    const f3 cp = p3i_mid(cpi);
    const f3 porg = ctx_offset_origin(cpi, cn, sub3(pc, cp));
    if (sqdist3(porg, pc) <= sqrf(s.radius)) {
        sphere_sample_area(s, u0, u1, q);
        const f3 sp = p3i_mid(q->pi);
        f3 wi = sub3(sp, cp);
        if (sqlen3(wi) == 0) return false;
        wi = norm3(wi);
        q->pdf /= absdot3(q->n, neg3(wi)) / sqdist3(cp, sp);
        return !isinf(q->pdf);
    }
    const float sin_max = s.radius / sqrtf(sqdist3(cp, pc));
    const float sin2_max = sqrf(sin_max);
    const float cos_max = safe_sqrtf(1 - sin2_max);
    float omc = 1 - cos_max;
    float cos_t = (cos_max - 1) * u0 + 1;
    float sin2_t = 1 - sqrf(cos_t);
    if (sin2_max < 0.00068523f) {  // sin^2(1.5 deg): Taylor expansion for small cones
        sin2_t = sin2_max * u0;
        cos_t = sqrtf(1 - sin2_t);
        omc = sin2_max / 2;
    }
    const float cos_a = sin2_t / sin_max + cos_t * safe_sqrtf(1 - sin2_t / sqrf(sin_max));
    const float sin_a = safe_sqrtf(1 - sqrf(cos_a));
    const float phi = u1 * 2 * kPi;
    const float sa = clampf(sin_a, -1, 1);
    const f3 w = {sa * tcos(phi), sa * tsin(phi), clampf(cos_a, -1, 1)};
    const f3 fz = norm3(sub3(pc, cp));
    f3 fx, fy;
    basis(fz, &fx, &fy);
    const f3 mw = neg3(w);
    f3 n = add3(add3(mul3(mw.x, fx), mul3(mw.y, fy)), mul3(mw.z, fz));
    const f3 p = add3(pc, mul3(s.radius, n));
    if (s.reverse) n = mul3(-1.f, n);
    q->pi = p3i_make(p, mul3(gamma_n(5), abs3(p)));
    q->n = n;
    q->pdf = 1 / (2 * kPi * omc);
    return true;
}

// Disk::sample(u) / sample(ctx, u) (shapes/disk.cu:39-94).
PBRTM_HD bool disk_sample(const DShape &s, const P3i &cpi, float u0, float u1, QSample *q) {
    float dx, dy;
    concentric_disk(u0, u1, &dx, &dy);
    q->pi = xf_point_iv(s.m, p3i_exact(f3{dx * s.radius, dy * s.radius, s.height}));
    f3 n = norm3(xf_normal(s.inv, f3{0.f, 0.f, 1.f}));
    if (s.reverse) n = mul3(-1.f, n);
    q->n = n;
    q->pdf = float(1.0 / double(disk_area(s)));
    const f3 cp = p3i_mid(cpi), sp = p3i_mid(q->pi);
    f3 wi = sub3(sp, cp);
    if (sqlen3(wi) == 0) return false;
    wi = norm3(wi);
    q->pdf /= absdot3(q->n, neg3(wi)) / sqdist3(cp, sp);
    return !isinf(q->pdf);
}

// The solid-angle pdf of a direction by intersecting the shape (Sphere::pdf inside
// the sphere, Disk::pdf): (1 / area) / (|n . -wi| / d^2), in double, 0 if infinite.
PBRTM_HD float hit_area_pdf(const DShape &s, const P3i &cpi, f3 cn, f3 wi) {
    const f3 o = ctx_offset_origin(cpi, cn, wi);
    QHit h;
    const bool hit = s.type == kShapeSphere ? sphere_hit(s, o, wi, __builtin_inff(), &h)
                                            : disk_hit(s, o, wi, __builtin_inff(), &h);
    if (!hit) return 0.f;
    const Surf g = s.type == kShapeSphere ? sphere_surf(s, h.p) : disk_surf(s, h.p);
    const float area = s.type == kShapeSphere ? sphere_area(s) : disk_area(s);
    const float pdf = float((1.0 / double(area)) /
                            double(absdot3(g.n, neg3(wi)) / sqdist3(p3i_mid(cpi), p3i_mid(g.pi))));
    return isinf(pdf) ? 0.f : pdf;
}

// Sphere::pdf (shapes/sphere.cu:257-297), Disk::pdf (shapes/disk.cu:20-37).
PBRTM_HD float shape_light_pdf(const DShape &s, const P3i &cpi, f3 cn, f3 wi) {
    if (s.type == kShapeSphere) {
        const f3 pc = xf_point(s.m, f3{0.f, 0.f, 0.f});
        const f3 cp = p3i_mid(cpi);
        const f3 porg = ctx_offset_origin(cpi, cn, sub3(pc, cp));
        if (sqdist3(porg, pc) > sqrf(s.radius)) {
            const float sin2_max = s.radius * s.radius / sqdist3(cp, pc);
            const float cos_max = safe_sqrtf(1 - sin2_max);
            float omc = 1 - cos_max;
            if (sin2_max < 0.00068523f) omc = sin2_max / 2;
            return float(1.0 / (2.0 * double(kPi) * double(omc)));
        }
    }
    return hit_area_pdf(s, cpi, cn, wi);
}

PBRTM_HD bool shape_light_sample(const DShape &s, const P3i &cpi, f3 cn, float u0, float u1, QSample *q) {
    return s.type == kShapeSphere ? sphere_sample(s, cpi, cn, u0, u1, q) : disk_sample(s, cpi, u0, u1, q);
}

}  // namespace pt
}  // namespace pbrtm
