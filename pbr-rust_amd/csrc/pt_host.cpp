// Host-side scene set-up for the path tracer: render-space geometry, the
// spectral globals (CIE, normalised D65, sRGB colour space, metal spectra), the
// Jakob-Hanika RGB-to-spectrum coefficients, the perspective camera matrices,
// the tabulated Mitchell filter and the power light sampler. Everything here is
// FP32 with the reference's double promotions and FMA placements; it runs once
// per scene.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../data/spectra_data.h"
#include "../../include/pbrtm.h"
#include "pathtrace.h"

namespace pbrtm {
namespace pt {
namespace {

constexpr float kPiF = 3.14159274101257324f;  // acos(-1) rounded to float
constexpr int kLMin = 360, kLMax = 830;

float lerp_f(float t, float a, float b) { return (1 - t) * a + t * b; }
float dop_h(float a, float b, float c, float d) {
    const float cd = c * d;
    return std::fma(a, b, -cd) + std::fma(-c, d, cd);
}
float sop_h(float a, float b, float c, float d) {
    const float cd = c * d;
    return std::fma(a, b, cd) + std::fma(c, d, -cd);
}

struct V {
    float x, y, z;
};
V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float vdot(V a, V b) { return std::fma(a.x, b.x, sop_h(a.y, b.y, a.z, b.z)); }
V vcross(V a, V b) { return {dop_h(a.y, b.z, a.z, b.y), dop_h(a.z, b.x, a.x, b.z), dop_h(a.x, b.y, a.y, b.x)}; }
float vlen(V a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
V vnorm(V a) {
    const float l = vlen(a);
    return {a.x / l, a.y / l, a.z / l};
}

// Compensated dot product of pairs (util/compensated_float.h), folded from the last pair.
float comp_dot(const float *a, const float *b, int n) {
    float v = a[n - 1] * b[n - 1];
    float err = std::fma(a[n - 1], b[n - 1], -v);
    for (int i = n - 2; i >= 0; --i) {
        const float ab = a[i] * b[i];
        const float ab_err = std::fma(a[i], b[i], -ab);
        const float s = ab + v;
        const float d = s - ab;
        err = ab_err + (err + ((ab - (s - d)) + (v - d)));
        v = s;
    }
    return v + err;
}

// ------------------------------------------------------------ matrices
struct Mat4 {
    float m[4][4];
};
Mat4 mat4_diag(float a, float b, float c) {
    Mat4 r{};
    r.m[0][0] = a, r.m[1][1] = b, r.m[2][2] = c, r.m[3][3] = 1;
    return r;
}
Mat4 mat4_mul(const Mat4 &a, const Mat4 &b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = 0;
            for (int k = 0; k < 4; ++k) acc = std::fma(a.m[i][k], b.m[k][j], acc);
            r.m[i][j] = acc;
        }
    return r;
}
float cdot3(float a0, float b0, float a1, float b1, float a2, float b2) {
    const float a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2};
    return comp_dot(a, b, 3);
}
// Inverse by Laplace expansion over 2x2 minors (euclidean_space/squared_matrix.cu).
Mat4 mat4_inv(const Mat4 &A) {
    const auto &m = A.m;
    const float s0 = dop_h(m[0][0], m[1][1], m[1][0], m[0][1]), s1 = dop_h(m[0][0], m[1][2], m[1][0], m[0][2]);
    const float s2 = dop_h(m[0][0], m[1][3], m[1][0], m[0][3]), s3 = dop_h(m[0][1], m[1][2], m[1][1], m[0][2]);
    const float s4 = dop_h(m[0][1], m[1][3], m[1][1], m[0][3]), s5 = dop_h(m[0][2], m[1][3], m[1][2], m[0][3]);
    const float c0 = dop_h(m[2][0], m[3][1], m[3][0], m[2][1]), c1 = dop_h(m[2][0], m[3][2], m[3][0], m[2][2]);
    const float c2 = dop_h(m[2][0], m[3][3], m[3][0], m[2][3]), c3 = dop_h(m[2][1], m[3][2], m[3][1], m[2][2]);
    const float c4 = dop_h(m[2][1], m[3][3], m[3][1], m[2][3]), c5 = dop_h(m[2][2], m[3][3], m[3][2], m[2][3]);
    const float da[6] = {s0, -s1, s2, s3, s5, -s4}, db[6] = {c5, c4, c3, c2, c0, c1};
    const float det = comp_dot(da, db, 6);
    if (det == 0) throw std::runtime_error("singular camera matrix");
    const float s = float(1.0 / double(det));
    Mat4 r;
    r.m[0][0] = s * cdot3(m[1][1], c5, m[1][3], c3, -m[1][2], c4);
    r.m[0][1] = s * cdot3(-m[0][1], c5, m[0][2], c4, -m[0][3], c3);
    r.m[0][2] = s * cdot3(m[3][1], s5, m[3][3], s3, -m[3][2], s4);
    r.m[0][3] = s * cdot3(-m[2][1], s5, m[2][2], s4, -m[2][3], s3);
    r.m[1][0] = s * cdot3(-m[1][0], c5, m[1][2], c2, -m[1][3], c1);
    r.m[1][1] = s * cdot3(m[0][0], c5, m[0][3], c1, -m[0][2], c2);
    r.m[1][2] = s * cdot3(-m[3][0], s5, m[3][2], s2, -m[3][3], s1);
    r.m[1][3] = s * cdot3(m[2][0], s5, m[2][3], s1, -m[2][2], s2);
    r.m[2][0] = s * cdot3(m[1][0], c4, m[1][3], c0, -m[1][1], c2);
    r.m[2][1] = s * cdot3(-m[0][0], c4, m[0][1], c2, -m[0][3], c0);
    r.m[2][2] = s * cdot3(m[3][0], s4, m[3][3], s0, -m[3][1], s2);
    r.m[2][3] = s * cdot3(-m[2][0], s4, m[2][1], s2, -m[2][3], s0);
    r.m[3][0] = s * cdot3(-m[1][0], c3, m[1][1], c1, -m[1][2], c0);
    r.m[3][1] = s * cdot3(m[0][0], c3, m[0][2], c0, -m[0][1], c1);
    r.m[3][2] = s * cdot3(-m[3][0], s3, m[3][1], s1, -m[3][2], s0);
    r.m[3][3] = s * cdot3(m[2][0], s3, m[2][2], s0, -m[2][1], s1);
    return r;
}
struct Xform {  // a matrix and its inverse, as Transform keeps them
    Mat4 m, inv;
};
Xform xmul(const Xform &a, const Xform &b) { return {mat4_mul(a.m, b.m), mat4_mul(b.inv, a.inv)}; }
Xform xinv(const Xform &a) { return {a.inv, a.m}; }
Xform xscale(float x, float y, float z) { return {mat4_diag(x, y, z), mat4_diag(float(1.0) / x, float(1.0) / y, float(1.0) / z)}; }
Xform xtranslate(float x, float y, float z) {
    Xform t{mat4_diag(1, 1, 1), mat4_diag(1, 1, 1)};
    t.m.m[0][3] = x, t.m.m[1][3] = y, t.m.m[2][3] = z;
    t.inv.m[0][3] = -x, t.inv.m[1][3] = -y, t.inv.m[2][3] = -z;
    return t;
}
V xpoint(const Mat4 &M, V p) {
    const auto &m = M.m;
    const float xp = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    const float yp = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    const float zp = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    const float wp = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (wp == 1.0f) return {xp, yp, zp};
    return {xp / wp, yp / wp, zp / wp};
}

struct Mat3 {
    float m[3][3];
};
Mat3 mat3_mul(const Mat3 &a, const Mat3 &b) {
    Mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float acc = 0;
            for (int k = 0; k < 3; ++k) acc = std::fma(a.m[i][k], b.m[k][j], acc);
            r.m[i][j] = acc;
        }
    return r;
}
Mat3 mat3_inv(const Mat3 &a) {
    const auto &m = a.m;
    const float det = std::fma(m[0][2], dop_h(m[1][0], m[2][1], m[1][1], m[2][0]),
                               dop_h(m[0][0], dop_h(m[1][1], m[2][2], m[1][2], m[2][1]), m[0][1],
                                     dop_h(m[1][0], m[2][2], m[1][2], m[2][0])));
    if (det == 0) throw std::runtime_error("singular colour matrix");
    const float id = float(1.0 / double(det));
    Mat3 r;
    r.m[0][0] = id * dop_h(m[1][1], m[2][2], m[1][2], m[2][1]);
    r.m[1][0] = id * dop_h(m[1][2], m[2][0], m[1][0], m[2][2]);
    r.m[2][0] = id * dop_h(m[1][0], m[2][1], m[1][1], m[2][0]);
    r.m[0][1] = id * dop_h(m[0][2], m[2][1], m[0][1], m[2][2]);
    r.m[1][1] = id * dop_h(m[0][0], m[2][2], m[0][2], m[2][0]);
    r.m[2][1] = id * dop_h(m[0][1], m[2][0], m[0][0], m[2][1]);
    r.m[0][2] = id * dop_h(m[0][1], m[1][2], m[0][2], m[1][1]);
    r.m[1][2] = id * dop_h(m[0][2], m[1][0], m[0][0], m[1][2]);
    r.m[2][2] = id * dop_h(m[0][0], m[1][1], m[0][1], m[1][0]);
    return r;
}

// ------------------------------------------------------------ spectra
struct PL {  // piecewise-linear spectrum
    std::vector<float> lam, val;
    float at(float l) const {
        const size_t n = lam.size();
        if (n == 0 || l < lam[0] || l > lam[n - 1]) return 0;
        long size = (long)n - 2, first = 1;
        while (size > 0) {
            const long half = size >> 1, mid = first + half;
            const bool r = lam[mid] <= l;
            first = r ? mid + 1 : first;
            size = r ? size - (half + 1) : half;
        }
        long o = std::min<long>(std::max<long>(first - 1, 0), (long)n - 2);
        const float t = (l - lam[o]) / (lam[o + 1] - lam[o]);
        return lerp_f(t, val[o], val[o + 1]);
    }
    float dot(const PL &s) const {  // FP32 sum over integer nanometres
        float sum = 0;
        for (int l = kLMin; l <= kLMax; ++l) sum += at(float(l)) * s.at(float(l));
        return sum;
    }
};
PL interleaved(const float *s, int len) {
    PL p;
    if (s[0] > kLMin) {
        p.lam.push_back(kLMin - 1);
        p.val.push_back(s[1]);
    }
    for (int i = 0; i < len / 2; ++i) {
        p.lam.push_back(s[2 * i]);
        p.val.push_back(s[2 * i + 1]);
    }
    if (p.lam.back() < kLMax) {
        p.lam.push_back(kLMax + 1);
        p.val.push_back(p.val.back());
    }
    return p;
}

// Jakob-Hanika optimisation (rgb_to_spectrum_data.cu), column by column.
class SpectrumTable {
  public:
    static constexpr int kRes = 64, kSamples = 95, kFine = (kSamples - 1) * 3 + 1;
    SpectrumTable() {
        static const double to_rgb[3][3] = {
            {3.240479, -1.537150, -0.498535}, {-0.969256, 1.875991, 0.041556}, {0.055648, -0.204043, 1.057311}};
        std::memset(tbl_, 0, sizeof(tbl_));
        std::memset(white_, 0, sizeof(white_));
        const double h = (830.0 - 360.0) / double(kFine - 1);
        for (int i = 0; i < kFine; ++i) {
            const double l = 360.0 + i * h;
            const double xyz[3] = {interp(pbrtm_data::kCieX5nm, l), interp(pbrtm_data::kCieY5nm, l),
                                   interp(pbrtm_data::kCieZ5nm, l)};
            const double I = interp(pbrtm_data::kD65_5nm, l);
            const double ratio = (i == 0 || i == kFine - 1) ? 1.0 : (((i - 1) % 3 == 2) ? 2.0 : 3.0);
            const double w = 3.0 / 8.0 * h * ratio;
            lam_[i] = l;
            for (int k = 0; k < 3; ++k)
                for (int j = 0; j < 3; ++j) tbl_[k][i] += to_rgb[k][j] * xyz[j] * I * w;
            for (int k = 0; k < 3; ++k) white_[k] += xyz[k] * I * w;
        }
        for (int k = 0; k < kRes; ++k) {
            const double x = double(k) / double(kRes - 1);
            const double s1 = x * x * (3.0 - 2.0 * x);
            z_[k] = s1 * s1 * (3.0 - 2.0 * s1);
        }
    }
    // Sigmoid coefficients of an RGB in [0, inf)^3 (RGBtoSpectrumTable::operator()).
    std::array<float, 3> lookup(const float rgb[3]) {
        if (rgb[0] == rgb[1] && rgb[1] == rgb[2])
            return {0.f, 0.f, float(double(rgb[0] - 0.5f) / std::sqrt(double(rgb[0]) * (1.0 - double(rgb[0]))))};
        const int mc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0 : 2) : ((rgb[1] > rgb[2]) ? 1 : 2);
        const double z = rgb[mc];
        const double x = double(rgb[(mc + 1) % 3] * (kRes - 1)) / z;
        const double y = double(rgb[(mc + 2) % 3] * (kRes - 1)) / z;
        const int xi = std::min((int)x, kRes - 2), yi = std::min((int)y, kRes - 2);
        long size = kRes - 2, first = 1;
        while (size > 0) {
            const long half = size >> 1, mid = first + half;
            const bool r = z_[mid] < z;
            first = r ? mid + 1 : first;
            size = r ? size - (half + 1) : half;
        }
        const int zi = int(std::min<long>(std::max<long>(first - 1, 0), kRes - 2));
        const float dx = float(x - xi), dy = float(y - yi);
        const float dz = float((z - z_[zi]) / (z_[zi + 1] - z_[zi]));
        const Column &a = column(mc, yi, xi), &b = column(mc, yi, xi + 1);
        const Column &c = column(mc, yi + 1, xi), &d = column(mc, yi + 1, xi + 1);
        std::array<float, 3> out;
        for (int q = 0; q < 3; ++q) {
            auto f = [&](const Column &col, int dk) { return float(col[zi + dk][q]); };
            out[q] = lerp_f(dz, lerp_f(dy, lerp_f(dx, f(a, 0), f(b, 0)), lerp_f(dx, f(c, 0), f(d, 0))),
                            lerp_f(dy, lerp_f(dx, f(a, 1), f(b, 1)), lerp_f(dx, f(c, 1), f(d, 1))));
        }
        return out;
    }

    // J[q][k] = d c_q / d rgb_k of lookup(): the trilinear interpolation's
    // derivative times that of the cell coordinates x = rgb[(m+1)%3] (res-1) / z,
    // y = rgb[(m+2)%3] (res-1) / z, z = rgb[m] (grey inputs take this general branch).
    void jacobian(const float rgb[3], double J[3][3]) { jacobian_impl(rgb, J, true); }

    // out[m][k] = sum_q J_m[q][k] coef[m][q] for n RGBs, on all host threads (the
    // columns are computed first, so the threads only read the cache).
    void jacobian_transpose(const float *rgb, const float *coef, size_t n, float *out) {
        prefetch(rgb, n, true);
        const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 64u));
        std::vector<std::thread> ts;
        for (unsigned t = 0; t < nt; ++t)
            ts.emplace_back([&, t] {
                for (size_t m = t; m < n; m += nt) {
                    double J[3][3];
                    jacobian_impl(rgb + 3 * m, J, false);
                    for (int k = 0; k < 3; ++k) {
                        double v = 0;
                        for (int q = 0; q < 3; ++q) v += J[q][k] * double(coef[3 * m + q]);
                        out[3 * m + k] = float(v);
                    }
                }
            });
        for (auto &th : ts) th.join();
    }

  private:
    void jacobian_impl(const float rgb[3], double J[3][3], bool locked) {
        const int mc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0 : 2) : ((rgb[1] > rgb[2]) ? 1 : 2);
        const double z = rgb[mc];
        if (!(z > 0) || rgb[0] < 0 || rgb[1] < 0 || rgb[2] < 0) {  // black: no general-branch cell
            for (int q = 0; q < 9; ++q) J[q / 3][q % 3] = 0;
            return;
        }
        const double x = double(rgb[(mc + 1) % 3] * (kRes - 1)) / z;
        const double y = double(rgb[(mc + 2) % 3] * (kRes - 1)) / z;
        const int xi = std::min((int)x, kRes - 2), yi = std::min((int)y, kRes - 2);
        int zi = 0;
        while (zi < kRes - 2 && z_[zi + 1] < z) ++zi;
        const double dx = x - xi, dy = y - yi, h = z_[zi + 1] - z_[zi], dz = (z - z_[zi]) / h;
        auto get = [&](int l, int j, int i) -> const Column & {
            return locked ? column(l, j, i) : cols_.find((l * kRes + j) * kRes + i)->second;
        };
        const Column &a = get(mc, yi, xi), &b = get(mc, yi, xi + 1);
        const Column &c = get(mc, yi + 1, xi), &d = get(mc, yi + 1, xi + 1);
        auto mix = [](double t, double u, double v) { return u + t * (v - u); };
        for (int q = 0; q < 3; ++q) {
            auto f = [&](const Column &col, int dk) { return double(float(col[zi + dk][q])); };
            double gx[2], gy[2], v[2];
            for (int k = 0; k < 2; ++k) {
                gx[k] = mix(dy, f(b, k) - f(a, k), f(d, k) - f(c, k));
                gy[k] = mix(dx, f(c, k) - f(a, k), f(d, k) - f(b, k));
                v[k] = mix(dy, mix(dx, f(a, k), f(b, k)), mix(dx, f(c, k), f(d, k)));
            }
            const double ddx = mix(dz, gx[0], gx[1]), ddy = mix(dz, gy[0], gy[1]), ddz = v[1] - v[0];
            J[q][mc] = -ddx * x / z - ddy * y / z + ddz / h;
            J[q][(mc + 1) % 3] = ddx * (kRes - 1) / z;
            J[q][(mc + 2) % 3] = ddy * (kRes - 1) / z;
        }
    }

  public:

  private:
    using Column = std::array<std::array<double, 3>, kRes>;
    static double interp(const double *d, double x) {
        x -= 360.0;
        x *= (kSamples - 1) / (830.0 - 360.0);
        int o = (int)x;
        o = std::max(0, std::min(o, kSamples - 2));
        const double w = x - o;
        return (1.0 - w) * d[o] + w * d[o + 1];
    }
    void lab(double *p) const {
        static const double to_xyz[3][3] = {
            {0.412453, 0.357580, 0.180423}, {0.212671, 0.715160, 0.072169}, {0.019334, 0.119193, 0.950227}};
        double X = 0, Y = 0, Z = 0;
        for (int j = 0; j < 3; ++j) {
            X += p[j] * to_xyz[0][j];
            Y += p[j] * to_xyz[1][j];
            Z += p[j] * to_xyz[2][j];
        }
        auto f = [](double t) {
            const double d = 6.0 / 29.0;
            return t > d * d * d ? std::cbrt(t) : t / (d * d * 3.0) + (4.0 / 29.0);
        };
        p[0] = 116.0 * f(Y / white_[1]) - 16.0;
        p[1] = 500.0 * (f(X / white_[0]) - f(Y / white_[1]));
        p[2] = 200.0 * (f(Y / white_[1]) - f(Z / white_[2]));
    }
    void residual(const double *c, const double *rgb, double *res) const {
        double out[3] = {0, 0, 0};
        for (int i = 0; i < kFine; ++i) {
            const double l = (lam_[i] - 360.0) / (830.0 - 360.0);
            double x = 0.0;
            for (int k = 0; k < 3; ++k) x = x * l + c[k];
            const double s = 0.5 * x / std::sqrt(1.0 + x * x) + 0.5;
            for (int j = 0; j < 3; ++j) out[j] += tbl_[j][i] * s;
        }
        lab(out);
        std::memcpy(res, rgb, sizeof(double) * 3);
        lab(res);
        for (int j = 0; j < 3; ++j) res[j] -= out[j];
    }
    void newton(const double rgb[3], double c[3]) const {
        for (int it = 0; it < 15; ++it) {
            double rows[3][3], *J[3] = {rows[0], rows[1], rows[2]};
            double r[3], r0[3], r1[3], t[3];
            residual(c, rgb, r);
            for (int i = 0; i < 3; ++i) {
                std::memcpy(t, c, sizeof(t));
                t[i] -= 1e-4;
                residual(t, rgb, r0);
                std::memcpy(t, c, sizeof(t));
                t[i] += 1e-4;
                residual(t, rgb, r1);
                for (int j = 0; j < 3; ++j) J[j][i] = (r1[j] - r0[j]) * 1.0 / (2 * 1e-4);
            }
            // LU with partial pivoting by row swaps, then forward/back substitution.
            int P[4] = {0, 1, 2, 3};
            for (int i = 0; i < 3; ++i) {
                double mx = 0.0;
                int im = i;
                for (int k = i; k < 3; ++k)
                    if (std::fabs(J[k][i]) > mx) mx = std::fabs(J[k][i]), im = k;
                if (mx < 1e-15) throw std::runtime_error("RGB to spectrum: singular Jacobian");
                if (im != i) {
                    std::swap(P[i], P[im]);
                    std::swap(J[i], J[im]);
                    P[3]++;
                }
                for (int j = i + 1; j < 3; ++j) {
                    J[j][i] /= J[i][i];
                    for (int k = i + 1; k < 3; ++k) J[j][k] -= J[j][i] * J[i][k];
                }
            }
            double x[3];
            for (int i = 0; i < 3; ++i) {
                x[i] = r[P[i]];
                for (int k = 0; k < i; ++k) x[i] -= J[i][k] * x[k];
            }
            for (int i = 2; i >= 0; --i) {
                for (int k = i + 1; k < 3; ++k) x[i] -= J[i][k] * x[k];
                x[i] = x[i] / J[i][i];
            }
            double rr = 0.0;
            for (int j = 0; j < 3; ++j) {
                c[j] -= x[j];
                rr += r[j] * r[j];
            }
            const double m = std::max(std::max(c[0], c[1]), c[2]);
            if (m > 200)
                for (int j = 0; j < 3; ++j) c[j] *= 200 / m;
            if (rr < 1e-6) break;
        }
    }
  public:
    // Computes, on all host threads, the table columns that lookups of these RGBs
    // ([n][3]) will read and that are not cached yet (the reference fills the
    // whole table on a thread pool at start-up; here only what a scene needs).
    void prefetch(const float *rgb, size_t n, bool include_grey = false) {
        std::vector<int> keys;
        {
            std::lock_guard<std::mutex> g(mu_);
            std::vector<char> want(3 * kRes * kRes, 0);
            for (size_t k = 0; k < n; ++k) {
                const float *c = rgb + 3 * k;
                if (!include_grey && c[0] == c[1] && c[1] == c[2]) continue;
                const int mc = (c[0] > c[1]) ? ((c[0] > c[2]) ? 0 : 2) : ((c[1] > c[2]) ? 1 : 2);
                const double z = c[mc];
                if (!(z > 0)) continue;
                const double x = double(c[(mc + 1) % 3] * (kRes - 1)) / z;
                const double y = double(c[(mc + 2) % 3] * (kRes - 1)) / z;
                const int xi = std::min((int)x, kRes - 2), yi = std::min((int)y, kRes - 2);
                if (xi < 0 || yi < 0) continue;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) want[(mc * kRes + yi + dy) * kRes + xi + dx] = 1;
            }
            for (int key = 0; key < int(want.size()); ++key)
                if (want[key] && !cols_.count(key)) keys.push_back(key);
        }
        if (keys.empty()) return;
        std::vector<Column> out(keys.size());
        const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 64u));
        std::vector<std::thread> ts;
        for (unsigned t = 0; t < nt; ++t)
            ts.emplace_back([&, t] {
                for (size_t k = t; k < keys.size(); k += nt)
                    out[k] = compute(keys[k] / (kRes * kRes), (keys[k] / kRes) % kRes, keys[k] % kRes);
            });
        for (auto &th : ts) th.join();
        std::lock_guard<std::mutex> g(mu_);
        for (size_t k = 0; k < keys.size(); ++k) cols_.emplace(keys[k], out[k]);
    }

  private:
    const Column &column(int l, int j, int i) {
        std::lock_guard<std::mutex> g(mu_);
        const int key = (l * kRes + j) * kRes + i;
        auto it = cols_.find(key);
        if (it != cols_.end()) return it->second;
        return cols_.emplace(key, compute(l, j, i)).first->second;
    }

    // Column (l, j, i) of the table: Gauss-Newton fits along z, warm-started from
    // z index kRes / 5 upwards, then from there downwards.
    Column compute(int l, int j, int i) const {
        Column col;
        const double y = j / double(kRes - 1), x = i / double(kRes - 1);
        const double c0 = 360.0, c1 = 1.0 / (830.0 - 360.0);
        const int start = kRes / 5;
        auto put = [&](int k, const double *c) {
            col[k][0] = c[0] * (c1 * c1);
            col[k][1] = c[1] * c1 - 2 * c[0] * c0 * (c1 * c1);
            col[k][2] = c[2] - c[1] * c0 * c1 + c[0] * ((c0 * c1) * (c0 * c1));
        };
        for (int pass = 0; pass < 2; ++pass) {
            double c[3] = {0, 0, 0}, rgb[3];
            for (int k = start; pass == 0 ? k < kRes : k >= 0; k += pass == 0 ? 1 : -1) {
                rgb[l] = z_[k];
                rgb[(l + 1) % 3] = x * z_[k];
                rgb[(l + 2) % 3] = y * z_[k];
                newton(rgb, c);
                put(k, c);
            }
        }
        return col;
    }

    double lam_[kFine], tbl_[3][kFine], white_[3], z_[kRes];
    std::map<int, Column> cols_;
    std::mutex mu_;
};

struct Globals {
    PL cie[3], d65, metals[8], glass_bk7;
    Mat3 xyz_from_rgb, rgb_from_xyz;
    float white[2];
    SpectrumTable table;
    Globals() {
        const float *c[3] = {pbrtm_data::kCieX1nm, pbrtm_data::kCieY1nm, pbrtm_data::kCieZ1nm};
        for (int k = 0; k < 3; ++k) {
            // 471 leading (0, 0) samples, as global_spectra.cu:10-21 builds its vectors
            cie[k].lam.assign(471, 0.f);
            cie[k].val.assign(471, 0.f);
            for (int i = 0; i < 471; ++i) {
                cie[k].lam.push_back(float(kLMin + i));
                cie[k].val.push_back(c[k][i]);
            }
        }
        d65 = interleaved(pbrtm_data::kD65Interleaved, pbrtm_data::kD65InterleavedLen);
        const float s = pbrtm_data::kCieYIntegral / d65.dot(cie[1]);
        for (float &v : d65.val) v *= s;
        const float *md[8] = {pbrtm_data::kAgEta, pbrtm_data::kAgK, pbrtm_data::kAlEta, pbrtm_data::kAlK,
                              pbrtm_data::kAuEta, pbrtm_data::kAuK, pbrtm_data::kCuEta, pbrtm_data::kCuK};
        const int ml[8] = {pbrtm_data::kAgEtaLen, pbrtm_data::kAgKLen, pbrtm_data::kAlEtaLen, pbrtm_data::kAlKLen,
                           pbrtm_data::kAuEtaLen, pbrtm_data::kAuKLen, pbrtm_data::kCuEtaLen, pbrtm_data::kCuKLen};
        for (int i = 0; i < 8; ++i) metals[i] = interleaved(md[i], ml[i]);
        glass_bk7 = interleaved(pbrtm_data::kGlassBK7Eta, pbrtm_data::kGlassBK7EtaLen);
        float w[3];
        for (int k = 0; k < 3; ++k) w[k] = d65.dot(cie[k]) / pbrtm_data::kCieYIntegral;
        white[0] = w[0] / (w[0] + w[1] + w[2]);
        white[1] = w[1] / (w[0] + w[1] + w[2]);
        const float xy[3][2] = {{0.64f, 0.33f}, {0.3f, 0.6f}, {0.15f, 0.06f}};
        Mat3 prim;
        for (int k = 0; k < 3; ++k) {
            prim.m[0][k] = xy[k][0] * 1.0f / xy[k][1];
            prim.m[1][k] = 1.0f;
            prim.m[2][k] = (1 - xy[k][0] - xy[k][1]) * 1.0f / xy[k][1];
        }
        const Mat3 pinv = mat3_inv(prim);
        Mat3 diag{};
        for (int k = 0; k < 3; ++k) diag.m[k][k] = comp_dot(pinv.m[k], w, 3);
        xyz_from_rgb = mat3_mul(prim, diag);
        rgb_from_xyz = mat3_inv(xyz_from_rgb);
    }
};
Globals &globals() {
    static Globals g;
    return g;
}

uint32_t pool_add(SceneHost &S, const PL &s) {
    const uint32_t off = uint32_t(S.spec_lam.size());
    S.spec_lam.insert(S.spec_lam.end(), s.lam.begin(), s.lam.end());
    S.spec_val.insert(S.spec_val.end(), s.val.begin(), s.val.end());
    return off;
}

// Piecewise-constant 1D distribution tables (util/piecewise_constant_1d.h:17-58).
float pc1d(const float *f, size_t n, float mn, float mx, float *func, float *cdf) {
    for (size_t i = 0; i < n; ++i) func[i] = std::fabs(f[i]);
    cdf[0] = 0;
    for (size_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] * (mx - mn) / float(n);
    const float integral = cdf[n];
    for (size_t i = 1; i < n + 1; ++i) cdf[i] = integral == 0 ? float(i) / float(n) : cdf[i] / integral;
    return integral;
}

}  // namespace

void rgb_to_sigmoid_jacobian_transpose(const float *rgb, const float *coef, size_t n, float *out) {
    globals().table.jacobian_transpose(rgb, coef, n, out);
}

void rgb_to_sigmoid_jacobian(const float rgb[3], double J[9]) {
    double m[3][3];
    globals().table.jacobian(rgb, m);
    for (int q = 0; q < 9; ++q) J[q] = m[q / 3][q % 3];
}

void rgb_to_sigmoid(const float rgb[3], float out[3]) {
    auto c = globals().table.lookup(rgb);
    out[0] = c[0], out[1] = c[1], out[2] = c[2];
}

bool mat4_inverse_ref(const float in[16], float out[16]) {
    Mat4 m;
    std::memcpy(m.m, in, sizeof(m.m));
    try {
        const Mat4 r = mat4_inv(m);
        std::memcpy(out, r.m, sizeof(r.m));
    } catch (const std::exception &) {
        return false;
    }
    return true;
}

void srgb_matrices(float a[9], float b[9], float w[2]) {
    const Globals &g = globals();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[3 * i + j] = g.xyz_from_rgb.m[i][j], b[3 * i + j] = g.rgb_from_xyz.m[i][j];
    w[0] = g.white[0], w[1] = g.white[1];
}

// SquareMatrix<3>::determinant of the upper 3x3 (squared_matrix.cu:8-17) < 0:
// Transform::swaps_handedness (transform.h:208-217).
bool swaps_handedness(const Mat4 &M) {
    const auto &m = M.m;
    const float minor12 = dop_h(m[1][1], m[2][2], m[1][2], m[2][1]);
    const float minor02 = dop_h(m[1][0], m[2][2], m[1][2], m[2][0]);
    const float minor01 = dop_h(m[1][0], m[2][1], m[1][1], m[2][0]);
    return std::fma(m[0][2], minor01, dop_h(m[0][0], minor12, m[0][1], minor02)) < 0;
}
bool affine(const Mat4 &M) { return M.m[3][0] == 0 && M.m[3][1] == 0 && M.m[3][2] == 0 && M.m[3][3] == 1; }
bool is_identity(const Xform &x) {  // Transform::is_identity (transform.h:219-239)
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            if (x.m.m[i][k] != x.inv.m[i][k] || x.m.m[i][k] != (i == k ? 1.f : 0.f)) return false;
    return true;
}
// Transform::operator()(Bounds3f) (transform.h:323-349).
void xbounds(const Mat4 &M, const float lo[3], const float hi[3], float olo[3], float ohi[3]) {
    const auto &m = M.m;
    for (int i = 0; i < 3; ++i) olo[i] = ohi[i] = m[i][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            const float a = m[i][k] * lo[k], b = m[i][k] * hi[k];
            olo[i] += std::min(a, b);
            ohi[i] += std::max(a, b);
        }
}

// Primitive i of the description is an analytic shape or an instanced triangle:
// its DShape, and in its triangle slot the box corners lo, hi, lo (the build and
// the aggregate's bounds see exactly Primitive::bounds). Instances whose
// render_from_instance is the identity stay plain triangles, as the reference adds
// the definition's primitives directly (scene_builder.cu:775-810).
template <typename CtmOf>
std::string add_shape(const pbrtm_shape &sh, const pbrtm_scene_desc &d, uint64_t i, const Xform &rw,
                      const Xform &wr, CtmOf ctm_of, SceneHost &S) {
    DShape q{};
    q.type = sh.type;
    const bool reverse = d.tri_reverse && d.tri_reverse[i];
    q.reverse = reverse ? 1u : 0u;
    Xform x;
    if (sh.type == PBRTM_SHAPE_SPHERE || sh.type == PBRTM_SHAPE_DISK) {
        x = xmul(rw, ctm_of(sh.xform));  // render_from_object = render_from_world * CTM
    } else if (sh.type == PBRTM_SHAPE_INSTANCE) {
        if (d.tri_light && d.tri_light[i] >= 0) return "area lights are not supported with object instancing";
        x = xmul(xmul(rw, ctm_of(sh.xform)), wr);  // render_from_instance (scene_builder.cu:785-786)
        if (is_identity(x)) return "";
    } else {
        return "unknown shape type";
    }
    if (!affine(x.m) || !affine(x.inv)) return "projective shape transforms are not supported";
    std::memcpy(q.m, x.m.m, sizeof(q.m));
    std::memcpy(q.inv, x.inv.m, sizeof(q.inv));
    float blo[3], bhi[3];
    const float phimax = float(double(kPiF) / 180.0 * double(std::min(std::max(sh.phimax, 0.f), 360.f)));
    if (sh.type == PBRTM_SHAPE_SPHERE) {  // Sphere::init (shapes/sphere.cu:24-40)
        const float r = sh.radius;
        q.radius = r;
        q.zmin = std::min(std::max(std::min(sh.zmin, sh.zmax), -r), r);
        q.zmax = std::min(std::max(std::max(sh.zmin, sh.zmax), -r), r);
        q.theta_zmin = float(std::acos(double(std::min(std::max(q.zmin / r, -1.f), 1.f))));
        q.theta_zmax = float(std::acos(double(std::min(std::max(q.zmax / r, -1.f), 1.f))));
        q.phimax = phimax;
        q.flip = q.reverse;  // Sphere::init leaves transform_swaps_handedness false
        const float lo[3] = {-r, -r, q.zmin}, hi[3] = {r, r, q.zmax};
        xbounds(x.m, lo, hi, blo, bhi);
    } else if (sh.type == PBRTM_SHAPE_DISK) {  // Disk::create (shapes/disk.cu:6-25)
        q.radius = sh.radius;
        q.height = sh.height;
        q.inner = sh.innerradius;
        q.phimax = phimax;
        q.flip = (reverse != swaps_handedness(x.m)) ? 1u : 0u;
        const float lo[3] = {-sh.radius, -sh.radius, sh.height}, hi[3] = {sh.radius, sh.radius, sh.height};
        xbounds(x.m, lo, hi, blo, bhi);
    } else {  // TransformedPrimitive::bounds (transformed_primitive.cu:10-13) of the triangle's box
        for (int k = 0; k < 9; ++k) q.p[k] = S.tri[9 * i + k];
        q.flip = q.reverse;
        float lo[3], hi[3];
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(std::min(q.p[c], q.p[3 + c]), q.p[6 + c]);
            hi[c] = std::max(std::max(q.p[c], q.p[3 + c]), q.p[6 + c]);
        }
        xbounds(x.m, lo, hi, blo, bhi);
    }
    for (int c = 0; c < 3; ++c) {
        S.tri[9 * i + c] = blo[c];
        S.tri[9 * i + 3 + c] = bhi[c];
        S.tri[9 * i + 6 + c] = blo[c];
    }
    if (S.prim_shape.empty()) S.prim_shape.assign(d.num_tris, 0u);
    S.shapes.push_back(q);
    S.prim_shape[i] = uint32_t(S.shapes.size());
    return "";
}

std::string build_scene_host(const void *vdesc, SceneHost &S) {
    const pbrtm_scene_desc &d = *static_cast<const pbrtm_scene_desc *>(vdesc);
    try {
        Globals &G = globals();
        if (d.width == 0 || d.height == 0) return "film resolution must be positive";
        if (d.num_tris > 0 && (!d.tri || !d.tri_material)) return "triangles need vertices and materials";
        if (d.num_materials > 0 && !d.materials) return "materials missing";
        if (d.num_tris >= (1ull << 31)) return "too many triangles";
        S.width = d.width;
        S.height = d.height;
        // spectra pool: CIE x, y, z, D65, then the metals on demand
        S.spec_lam.clear();
        S.spec_val.clear();
        for (int k = 0; k < 3; ++k) S.cie_off[k] = pool_add(S, G.cie[k]);
        S.cie_n = uint32_t(G.cie[0].lam.size());
        S.d65_off = pool_add(S, G.d65);
        S.d65_n = uint32_t(G.d65.lam.size());
        // the longest uniformly spaced tail of a grid: its first index and 1 / step
        auto uniform_tail = [](const std::vector<float> &lam, uint32_t *start, float *inv_step) {
            const size_t n = lam.size();
            *start = uint32_t(n ? n - 1 : 0);
            *inv_step = 0.f;
            if (n < 3) return;
            const float d = lam[n - 1] - lam[n - 2];
            if (!(d > 0)) return;
            size_t a = n - 2;
            while (a > 0 && lam[a] - lam[a - 1] == d) --a;
            *start = uint32_t(a);
            *inv_step = 1.0f / d;
        };
        uniform_tail(G.cie[0].lam, &S.cie_uni, &S.cie_inv_step);
        uniform_tail(G.d65.lam, &S.d65_uni, &S.d65_inv_step);
        S.cie_shared = (G.cie[0].lam == G.cie[1].lam && G.cie[0].lam == G.cie[2].lam) ? 1u : 0u;
        int metal_off[8], glass_off = -1;
        for (int i = 0; i < 8; ++i) metal_off[i] = -1;
        // materials (materials/diffuse_material.cu:22-29, conductor_material.cu:10-76)
        {
            std::vector<float> alb;
            for (uint32_t i = 0; i < d.num_materials; ++i)
                if (d.materials[i].type == PBRTM_MATERIAL_DIFFUSE)
                    for (int q = 0; q < 3; ++q) alb.push_back(std::max(0.f, d.materials[i].rgb[q]));
            G.table.prefetch(alb.data(), alb.size() / 3);
        }
        S.mats.clear();
        S.mat_rgb.assign(size_t(d.num_materials) * 3, 0.f);
        S.group_scale.assign(d.num_lights, 0.f);
        for (uint32_t g = 0; g < d.num_lights && d.lights; ++g) S.group_scale[g] = d.lights[g].scale;
        for (uint32_t i = 0; i < d.num_materials; ++i) {
            const pbrtm_material &m = d.materials[i];
            DMat dm{};
            dm.type = m.type;
            for (int q = 0; q < 3; ++q) S.mat_rgb[3 * size_t(i) + q] = std::max(0.f, m.rgb[q]);
            if (m.type == PBRTM_MATERIAL_DIFFUSE) {
                if (m.rgb[0] > 1 || m.rgb[1] > 1 || m.rgb[2] > 1) return "diffuse reflectance RGB above 1";
                const float c[3] = {std::max(0.f, m.rgb[0]), std::max(0.f, m.rgb[1]), std::max(0.f, m.rgb[2])};
                auto s = G.table.lookup(c);
                dm.c0 = s[0], dm.c1 = s[1], dm.c2 = s[2];
            } else if (m.type == PBRTM_MATERIAL_CONDUCTOR) {
                if (m.metal > 3) return "unknown metal";
                for (int q = 0; q < 2; ++q) {
                    const int id = 2 * int(m.metal) + q;
                    if (metal_off[id] < 0) metal_off[id] = int(pool_add(S, G.metals[id]));
                }
                dm.eta_off = uint32_t(metal_off[2 * m.metal]);
                dm.eta_n = uint32_t(G.metals[2 * m.metal].lam.size());
                dm.k_off = uint32_t(metal_off[2 * m.metal + 1]);
                dm.k_n = uint32_t(G.metals[2 * m.metal + 1].lam.size());
                float u = m.uroughness, v = m.vroughness;
                if (m.remap_roughness) u = std::sqrt(u), v = std::sqrt(v);
                dm.ax = u, dm.ay = v;
            } else if (m.type == PBRTM_MATERIAL_DIELECTRIC) {
                if (m.eta_named > 1) return "unknown named index of refraction";
                dm.eta_c = m.eta;
                if (m.eta_named == 1) {
                    if (glass_off < 0) glass_off = int(pool_add(S, G.glass_bk7));
                    dm.eta_off = uint32_t(glass_off);
                    dm.eta_n = uint32_t(G.glass_bk7.lam.size());
                }
                float u = m.uroughness, v = m.vroughness;
                if (m.remap_roughness) u = std::sqrt(u), v = std::sqrt(v);
                dm.ax = u, dm.ay = v;
            } else {
                return "unknown material type";
            }
            S.mats.push_back(dm);
        }
        // camera (base/camera.h:26-45 camera-world render space; cameras/perspective.cu:6-61)
        const V eye{d.eye[0], d.eye[1], d.eye[2]}, look{d.look[0], d.look[1], d.look[2]}, up{d.up[0], d.up[1], d.up[2]};
        Mat4 wfc{};
        wfc.m[0][3] = eye.x, wfc.m[1][3] = eye.y, wfc.m[2][3] = eye.z, wfc.m[3][3] = 1.0f;
        const V dir = vnorm(vsub(look, eye));
        if (vlen(vcross(vnorm(up), dir)) == 0.0f) return "LookAt: up is parallel to the view direction";
        const V right = vnorm(vcross(vnorm(up), dir));
        const V nup = vcross(dir, right);
        const V cols[3] = {right, nup, dir};
        for (int c = 0; c < 3; ++c) wfc.m[0][c] = cols[c].x, wfc.m[1][c] = cols[c].y, wfc.m[2][c] = cols[c].z;
        const Xform camera_from_world{mat4_inv(wfc), wfc};
        const Xform world_from_camera = xinv(camera_from_world);
        const V pc = xpoint(world_from_camera.m, {0, 0, 0});
        const Xform render_from_world = xinv(xtranslate(pc.x, pc.y, pc.z));
        const Xform render_from_camera = xmul(render_from_world, world_from_camera);
        const float aspect = float(d.width) / float(d.height);
        float sw[4];
        if (aspect > 1.0) sw[0] = -aspect, sw[1] = -1.0f, sw[2] = aspect, sw[3] = 1.0f;
        else sw[0] = -1.0f, sw[1] = float(-1.0 / aspect), sw[2] = 1.0f, sw[3] = float(1.0 / aspect);
        const Xform ndc_from_screen = xmul(xscale(float(1.0 / (sw[2] - sw[0])), float(1.0 / (sw[3] - sw[1])), 1.0f),
                                           xtranslate(-sw[0], -sw[3], 0.0f));
        const Xform raster_from_screen = xmul(xscale(float(d.width), float(-int(d.height)), 1.0f), ndc_from_screen);
        const Xform screen_from_raster = xinv(raster_from_screen);
        const float zn = 1e-2f, zf = 1000.0f;
        Mat4 persp{};
        persp.m[0][0] = 1, persp.m[1][1] = 1, persp.m[2][2] = zf / (zf - zn), persp.m[2][3] = -zf * zn / (zf - zn);
        persp.m[3][2] = 1;
        const float rad = float(double(kPiF) / 180.0 * double(d.fov));
        const double inv_tan = 1.0 / std::tan(double(rad) / 2.0);
        const Xform screen_from_camera = xmul(xscale(float(inv_tan), float(inv_tan), 1.0f), {persp, mat4_inv(persp)});
        const Xform camera_from_raster = xmul(xinv(screen_from_camera), screen_from_raster);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                S.cam_from_raster[4 * i + j] = camera_from_raster.m.m[i][j];
                S.render_from_camera[4 * i + j] = render_from_camera.m.m[i][j];
            }
        const V o = xpoint(render_from_camera.m, {0, 0, 0});
        S.cam_origin[0] = o.x, S.cam_origin[1] = o.y, S.cam_origin[2] = o.z;
        // geometry in render space and per-triangle lights (triangle_mesh.cu:39-57,
        // scene_builder.cu:620-630 one DiffuseAreaLight per emissive triangle)
        const uint64_t n = d.num_tris;
        // render_from_object = render_from_world * CTM (scene_builder.cu get_render_from_object)
        if (d.num_xforms > 0 && !d.xforms) return "transforms missing";
        std::vector<Mat4> rfo(d.num_xforms);
        for (uint32_t k = 0; k < d.num_xforms; ++k) {
            Mat4 x;
            std::memcpy(x.m, d.xforms + 16 * size_t(k), sizeof(x.m));
            rfo[k] = mat4_mul(render_from_world.m, x);
        }
        S.tri.resize(n * 9);
        S.tri_info.resize(n);
        S.lights.clear();
        S.shapes.clear();
        S.prim_shape.clear();
        if (d.tri_shape && d.num_shapes && !d.shapes) return "shapes missing";
        // the CTM's inverse as the front end composed it, else the reference's 4x4 inverse
        auto ctm_of = [&](uint32_t k) -> Xform {
            Xform x{mat4_diag(1, 1, 1), mat4_diag(1, 1, 1)};
            if (k == 0xFFFFFFFFu) return x;
            if (k >= d.num_xforms) throw std::runtime_error("shape transform index out of range");
            std::memcpy(x.m.m, d.xforms + 16 * size_t(k), sizeof(x.m.m));
            if (d.xforms_inv) std::memcpy(x.inv.m, d.xforms_inv + 16 * size_t(k), sizeof(x.inv.m));
            else x.inv = mat4_inv(x.m);
            return x;
        };
        const Xform world_from_render = xinv(render_from_world);
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t xk = d.tri_xform ? d.tri_xform[i] : 0xFFFFFFFFu;
            if (xk != 0xFFFFFFFFu && xk >= d.num_xforms) return "triangle transform index out of range";
            const Mat4 &rm = xk == 0xFFFFFFFFu ? render_from_world.m : rfo[xk];
            for (int v = 0; v < 3; ++v) {
                const V p = xpoint(rm, {d.tri[9 * i + 3 * v], d.tri[9 * i + 3 * v + 1], d.tri[9 * i + 3 * v + 2]});
                S.tri[9 * i + 3 * v] = p.x, S.tri[9 * i + 3 * v + 1] = p.y, S.tri[9 * i + 3 * v + 2] = p.z;
            }
            if (d.tri_material[i] >= d.num_materials) return "triangle material index out of range";
            const uint32_t sk = d.tri_shape ? d.tri_shape[i] : 0xFFFFFFFFu;
            if (sk != 0xFFFFFFFFu) {
                if (sk >= d.num_shapes) return "shape index out of range";
                const std::string e = add_shape(d.shapes[sk], d, i, render_from_world, world_from_render, ctm_of, S);
                if (!e.empty()) return e;
            }
            const bool flip = d.tri_reverse && d.tri_reverse[i];
            uint32_t light = 0;
            const int32_t li = d.tri_light ? d.tri_light[i] : -1;
            if (li >= 0) {
                if (uint32_t(li) >= d.num_lights || !d.lights) return "triangle light index out of range";
                const pbrtm_area_light &L = d.lights[li];
                DLight dl{};
                dl.tri = uint32_t(i);
                dl.two_sided = L.two_sided ? 1u : 0u;
                dl.rgb_scale = float(2.0 * double(std::max(std::max(L.rgb[0], L.rgb[1]), L.rgb[2])));
                float c[3] = {0, 0, 0};
                if (dl.rgb_scale > 0.0f)
                    for (int q = 0; q < 3; ++q) c[q] = std::max(0.f, L.rgb[q] / dl.rgb_scale);
                auto s = G.table.lookup(c);
                dl.c0 = s[0], dl.c1 = s[1], dl.c2 = s[2];
                dl.scale = L.scale / G.d65.dot(G.cie[1]);
                dl.group = uint32_t(li);
                S.lights.push_back(dl);
                light = uint32_t(S.lights.size());
            }
            S.tri_info[i] = make_uint2(d.tri_material[i], light | (flip ? 0x80000000u : 0u));
        }
        // Infinite lights join the list at their place in file order (the reference pushes
        // LightSource lights and per-triangle area lights as it parses, scene_builder.cu:494,
        // :637); UniformInfiniteLight::create (uniform_infinite_light.cu:4-30).
        S.num_inf = 0;
        S.inf_dist = 0.f;
        if (d.num_infinite_lights) {
            if (!d.infinite_lights) return "infinite lights missing";
            if (d.num_infinite_lights > kMaxInfinite) return "at most 4 infinite lights";
            std::vector<DLight> area = std::move(S.lights);
            S.lights.clear();
            std::vector<uint32_t> remap(area.size());
            for (size_t a = 0; a <= area.size(); ++a) {
                for (uint32_t q = 0; q < d.num_infinite_lights; ++q) {
                    const pbrtm_infinite_light &I = d.infinite_lights[q];
                    if (std::min<size_t>(I.light_index, area.size()) != a) continue;
                    DLight dl{};
                    dl.tri = kInfiniteLight;
                    dl.two_sided = 1u;
                    dl.group = 0xFFFFFFFFu;
                    float scale = I.scale / G.d65.dot(G.cie[1]);
                    if (I.illuminance > 0) scale *= I.illuminance / kPiF;
                    dl.scale = scale;
                    if (I.has_rgb) {
                        dl.rgb_scale = float(2.0 * double(std::max(std::max(I.rgb[0], I.rgb[1]), I.rgb[2])));
                        float c[3] = {0, 0, 0};
                        if (dl.rgb_scale > 0.0f)
                            for (int w = 0; w < 3; ++w) c[w] = std::max(0.f, I.rgb[w] / dl.rgb_scale);
                        auto k = G.table.lookup(c);
                        dl.c0 = k[0], dl.c1 = k[1], dl.c2 = k[2];
                    } else {  // the colour space illuminant itself: a sigmoid that is 1 everywhere
                        dl.rgb_scale = 1.f;
                        dl.c0 = dl.c1 = 0.f;
                        dl.c2 = std::numeric_limits<float>::infinity();
                    }
                    S.inf_idx[S.num_inf++] = uint32_t(S.lights.size());
                    S.lights.push_back(dl);
                }
                if (a < area.size()) {
                    remap[a] = uint32_t(S.lights.size());
                    S.lights.push_back(area[a]);
                }
            }
            for (uint64_t i = 0; i < n; ++i) {
                const uint32_t l = S.tri_info[i].y & 0x7FFFFFFFu;
                if (l) S.tri_info[i].y = (S.tri_info[i].y & 0x80000000u) | (remap[l - 1] + 1);
            }
            // Bounds3::bounding_sphere of the aggregate's bounds (uniform_infinite_light.h:17-19)
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (size_t i = 0; i < S.tri.size(); i += 3)
                for (int c = 0; c < 3; ++c) lo[c] = std::min(lo[c], S.tri[i + c]), hi[c] = std::max(hi[c], S.tri[i + c]);
            float ctr[3];
            bool inside = true;
            for (int c = 0; c < 3; ++c) {
                ctr[c] = (lo[c] + hi[c]) / 2;
                inside = inside && ctr[c] >= lo[c] && ctr[c] <= hi[c];
            }
            const float dx = ctr[0] - hi[0], dy = ctr[1] - hi[1], dz = ctr[2] - hi[2];
            const float radius = inside ? std::sqrt(dx * dx + dy * dy + dz * dz) : 0.0f;
            S.inf_dist = 2 * radius;
        }
        // power light sampler (power_light_sampler.cu:7-45, distribution_1d.h:46-70)
        S.light_pmf.clear();
        S.light_cdf.clear();
        if (!S.lights.empty()) {
            float lam[4], pdf[4];
            for (int i = 0; i < 4; ++i) {
                float up2 = 0.5f + float(i) / 4;
                if (up2 > 1) up2 -= 1;
                lam[i] = 538 - 138.888889f * float(std::atanh(double(0.85691062f - 1.82750197f * up2)));
                const float ch = float(std::cosh(double(0.0072f * (lam[i] - 538))));
                pdf[i] = (lam[i] < kLMin || lam[i] > kLMax) ? 0.0f : 0.0039398042f / (ch * ch);
            }
            std::vector<float> w;
            for (const DLight &l : S.lights) {
                const bool inf = l.tri == kInfiniteLight;
                float k = 0.f;
                if (!inf) {
                    const float *t = &S.tri[9 * size_t(l.tri)];
                    const V p0{t[0], t[1], t[2]}, p1{t[3], t[4], t[5]}, p2{t[6], t[7], t[8]};
                    float area = vlen(vcross(vsub(p1, p0), vsub(p2, p0))) * 0.5f;
                    const uint32_t sk = S.prim_shape.empty() ? 0u : S.prim_shape[l.tri];
                    if (sk) {  // Sphere::area, Disk::area (shapes/sphere.h:31-34, shapes/disk.h:19-22)
                        const DShape &q = S.shapes[sk - 1];
                        area = q.type == kShapeSphere ? sphere_area(q) : disk_area(q);
                    }
                    k = kPiF * float(l.two_sided ? 2 : 1) * area;
                } else {  // UniformInfiniteLight::phi: 4 pi^2 R^2 scale L(lambda)
                    const float R = S.inf_dist / 2;
                    k = 4 * kPiF * kPiF * (R * R) * l.scale;
                }
                float sum = 0;
                for (int i = 0; i < 4; ++i) {
                    const float poly = std::fma(lam[i], std::fma(lam[i], l.c0, l.c1), l.c2);
                    const float sg = std::isinf(poly) ? (poly > 0 ? 1.f : 0.f) : .5f + poly / (2 * std::sqrt(1 + poly * poly));
                    const float phi = inf ? (G.d65.at(lam[i]) * (l.rgb_scale * sg)) * k
                                          : ((G.d65.at(lam[i]) * (l.rgb_scale * sg)) * l.scale) * k;
                    sum += pdf[i] == 0.0f ? 0.0f : phi / pdf[i];
                }
                w.push_back(sum / 4);
            }
            double total = 0.0;
            for (float x : w) total += x;
            const size_t nl = w.size();
            for (size_t i = 0; i < nl; ++i) S.light_pmf.push_back(total == 0.0 ? float(1.0 / nl) : float(w[i] / total));
            S.light_cdf.push_back(S.light_pmf[0]);
            for (size_t i = 1; i < nl; ++i) S.light_cdf.push_back(S.light_cdf[i - 1] + S.light_pmf[i]);
        }
        // Mitchell filter radius 2, B = C = 1/3, tabulated at 32 samples per unit (filter.cu:146-161)
        const float r = 2.f, B = 1.f / 3.f, C = 1.f / 3.f;
        const uint32_t fn = uint32_t(32 * r);
        auto mitchell = [&](float x) {
            x = std::fabs(x);
            if (x <= 1) return ((12 - 9 * B - 6 * C) * x * x * x + (-18 + 12 * B + 6 * C) * x * x + (6 - 2 * B)) * (1.f / 6.f);
            if (x <= 2)
                return ((-B - 6 * C) * x * x * x + (6 * B + 30 * C) * x * x + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) *
                       (1.f / 6.f);
            return 0.f;
        };
        S.filt_n = fn;
        S.filt_radius = r;
        S.filt_f.resize(size_t(fn) * fn);
        for (uint32_t y = 0; y < fn; ++y)
            for (uint32_t x = 0; x < fn; ++x) {
                const float px = lerp_f((x + 0.5f) / fn, -r, r), py = lerp_f((y + 0.5f) / fn, -r, r);
                S.filt_f[x + fn * y] = mitchell(2 * px / r) * mitchell(2 * py / r);
            }
        S.filt_cond_func.resize(size_t(fn) * fn);
        S.filt_cond_cdf.resize(size_t(fn) * (fn + 1));
        S.filt_cond_int.resize(fn);
        for (uint32_t v = 0; v < fn; ++v)
            S.filt_cond_int[v] = pc1d(&S.filt_f[size_t(v) * fn], fn, -r, r, &S.filt_cond_func[size_t(v) * fn],
                                      &S.filt_cond_cdf[size_t(v) * (fn + 1)]);
        S.filt_marg_func.resize(fn);
        S.filt_marg_cdf.resize(fn + 1);
        S.filt_marg_int = pc1d(S.filt_cond_int.data(), fn, -r, r, S.filt_marg_func.data(), S.filt_marg_cdf.data());
        // output_rgb_from_sensor_rgb = rgb_from_xyz * identity (rgb_film.cu:67)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) S.out_from_sensor[3 * i + j] = G.rgb_from_xyz.m[i][j];
        // AmbientOcclusionIntegrator::create (ambient_occlusion.cu:9-23): the colour
        // space illuminant and 1 / its photometric integral
        S.ao_scale = 1.0f / G.d65.dot(G.cie[1]);
        // SurfaceNormalIntegrator::init (surface_normal.h:17-23): albedo spectra of
        // RGB (0.01, 0, 0), (0, 0.01, 0), (0, 0, 0.01)
        for (int c = 0; c < 3; ++c) {
            float rgb[3] = {0.f, 0.f, 0.f};
            rgb[c] = 0.01f;
            const auto k = G.table.lookup(rgb);
            for (int q = 0; q < 3; ++q) S.sn_c[3 * c + q] = k[q];
        }
    } catch (const std::exception &e) {
        return e.what();
    }
    return "";
}

}  // namespace pt
}  // namespace pbrtm
