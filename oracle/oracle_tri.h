// TEST INFRASTRUCTURE ONLY — part of the CPU oracle (see oracle_math.h).
//
// Triangle intersection and the hit-point quantities the ambient-occlusion
// bounce needs, restated from shapes/triangle.cu.
#pragma once

#include "oracle_math.h"

namespace orc {

struct TriHit {
    float b0, b1, b2, t;
};

// shapes/triangle.cu:217-328 — pbrt-v4 watertight ray/triangle test.
// Returns false for a miss; t_max is inclusive (tScaled == t_max*det hits).
inline bool intersect_triangle(V3 ro, V3 rd, float t_max, V3 p0, V3 p1, V3 p2, TriHit *out) {
    if (sqlen(cross(sub(p2, p0), sub(p1, p0))) == 0.0f) return false;

    V3 p0t = sub(p0, ro), p1t = sub(p1, ro), p2t = sub(p2, ro);
    int kz = max_component_index(vabs(rd));
    int kx = (kz + 1) % 3, ky = (kz + 2) % 3;
    V3 d = {rd[kx], rd[ky], rd[kz]};
    p0t = {p0t[kx], p0t[ky], p0t[kz]};
    p1t = {p1t[kx], p1t[ky], p1t[kz]};
    p2t = {p2t[kx], p2t[ky], p2t[kz]};

    float Sx = -d.x / d.z, Sy = -d.y / d.z, Sz = 1 / d.z;
    p0t.x += Sx * p0t.z; p0t.y += Sy * p0t.z;
    p1t.x += Sx * p1t.z; p1t.y += Sy * p1t.z;
    p2t.x += Sx * p2t.z; p2t.y += Sy * p2t.z;

    float e0 = dop(p1t.x, p2t.y, p1t.y, p2t.x);
    float e1 = dop(p2t.x, p0t.y, p2t.y, p0t.x);
    float e2 = dop(p0t.x, p1t.y, p0t.y, p1t.x);
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {  // FP64 re-evaluation on edges
        double p2txp1ty = double(p2t.x) * double(p1t.y);
        double p2typ1tx = double(p2t.y) * double(p1t.x);
        e0 = float(p2typ1tx - p2txp1ty);
        double p0txp2ty = double(p0t.x) * double(p2t.y);
        double p0typ2tx = double(p0t.y) * double(p2t.x);
        e1 = float(p0typ2tx - p0txp2ty);
        double p1txp0ty = double(p1t.x) * double(p0t.y);
        double p1typ0tx = double(p1t.y) * double(p0t.x);
        e2 = float(p1typ0tx - p1txp0ty);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0) return false;

    p0t.z *= Sz; p1t.z *= Sz; p2t.z *= Sz;
    float tScaled = e0 * p0t.z + e1 * p1t.z + e2 * p2t.z;
    if (det < 0 && (tScaled >= 0 || tScaled < t_max * det)) return false;
    if (det > 0 && (tScaled <= 0 || tScaled > t_max * det)) return false;

    float invDet = 1 / det;
    float b0 = e0 * invDet, b1 = e1 * invDet, b2 = e2 * invDet;
    float t = tScaled * invDet;

    float maxZt = max_component_value(vabs(V3{p0t.z, p1t.z, p2t.z}));
    float deltaZ = gammaf(3) * maxZt;
    float maxXt = max_component_value(vabs(V3{p0t.x, p1t.x, p2t.x}));
    float maxYt = max_component_value(vabs(V3{p0t.y, p1t.y, p2t.y}));
    float deltaX = gammaf(5) * (maxXt + maxZt);
    float deltaY = gammaf(5) * (maxYt + maxZt);
    float deltaE = 2 * (gammaf(2) * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
    float maxE = max_component_value(vabs(V3{e0, e1, e2}));
    float deltaT = 3 * (gammaf(3) * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * std::fabs(invDet);
    if (t <= deltaT) return false;
    *out = {b0, b1, b2, t};
    return true;
}

// shapes/triangle.cu:376-396 — hit point with its error bound, geometric normal.
struct HitGeom {
    P3fi pi;
    V3 n;
};
inline HitGeom hit_geometry(const TriHit &h, V3 p0, V3 p1, V3 p2) {
    V3 pHit = add(add(scale(h.b0, p0), scale(h.b1, p1)), scale(h.b2, p2));
    V3 pAbs = add(add(vabs(scale(h.b0, p0)), vabs(scale(h.b1, p1))), vabs(scale(h.b2, p2)));
    V3 pErr = scale(gammaf(7), pAbs);
    V3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
    return {make_p3fi(pHit, pErr), normalize(cross(dp02, dp12))};
}

// euclidean_space/vector3.h:160-166 — branchless orthonormal basis.
inline void coordinate_system(V3 v, V3 *a, V3 *b) {
    float sign = std::copysign(1.0, v.z);
    float aa = -1.0 / (sign + v.z);
    float bb = v.x * v.y * aa;
    *a = {1 + sign * (v.x * v.x) * aa, sign * bb, -sign * v.x};
    *b = {bb, sign + (v.y * v.y) * aa, -v.y};
}

// util/sampling.h:61-83 and :219-224 — concentric disk, cosine hemisphere.
inline V3 sample_cosine_hemisphere(float u0, float u1) {
    float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    float dx = 0.f, dy = 0.f;
    if (!(ox == 0.0f && oy == 0.0f)) {
        const float pi = float(std::acos(-1.0));
        float pi_over_4 = float(pi / 4.0);
        float r, theta;
        if (std::fabs(ox) > std::fabs(oy)) {
            r = ox;
            theta = pi_over_4 * (oy / ox);
        } else {
            r = oy;
            float pi_over_2 = float(pi / 2.0);
            theta = pi_over_2 - pi_over_4 * (ox / oy);
        }
        // transcendentals in double, rounded once to float (see pt_oracle.cpp)
        dx = r * float(std::cos(double(theta)));
        dy = r * float(std::sin(double(theta)));
    }
    float zz = float(1.0 - (dx * dx) - (dy * dy));
    float z = std::sqrt(std::max(0.0f, zz));
    return {dx, dy, z};
}

}  // namespace orc
