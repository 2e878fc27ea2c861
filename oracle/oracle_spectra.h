// TEST INFRASTRUCTURE ONLY — part of the CPU oracle (see oracle_math.h).
//
// Spectral set-up restated from the reference: piecewise-linear spectra
// (spectra/piecewise_linear_spectrum.cu:29-112), the CIE/D65 globals
// (spectrum_util/global_spectra.cu:7-62), the sRGB colour space
// (spectrum_util/rgb_color_space.h:12-44), the Jakob-Hanika RGB-to-spectrum
// optimisation (spectrum_util/rgb_to_spectrum_data.cu:1-427) and the sigmoid
// polynomial (spectrum_util/rgb_sigmoid_polynomial.h). The coefficient table is
// computed lazily, one (max component, y, x) column at a time, with the same
// warm-start chain the reference uses to fill the full table.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "../pbr-rust_amd/data/spectra_data.h"
#include "oracle_math.h"

namespace orc {

constexpr int kLambdaMin = 360, kLambdaMax = 830;
constexpr int kNS = 4;  // NSpectrumSamples

inline float lerpf(float t, float a, float b) { return (1 - t) * a + t * b; }  // basic_math.h:27-29

// basic_math.h:86-98 — largest index in [0, sz-2] whose predicate holds.
template <typename P>
inline size_t find_interval(size_t sz, const P &pred) {
    using ssize = std::make_signed_t<size_t>;
    ssize size = (ssize)sz - 2, first = 1;
    while (size > 0) {
        size_t half = (size_t)size >> 1, middle = first + half;
        bool r = pred(middle);
        first = r ? middle + 1 : first;
        size = r ? size - (half + 1) : half;
    }
    ssize v = first - 1;
    if (v < 0) v = 0;
    if (v > (ssize)sz - 2) v = (ssize)sz - 2;
    return (size_t)v;
}

struct PLSpectrum {  // piecewise_linear_spectrum.cu
    std::vector<float> lam, val;
    float operator()(float l) const {
        const size_t n = lam.size();
        if (n == 0 || l < lam[0] || l > lam[n - 1]) return 0;
        size_t o = find_interval(n, [&](size_t i) { return lam[i] <= l; });
        float t = (l - lam[o]) / (lam[o + 1] - lam[o]);
        return lerpf(t, val[o], val[o + 1]);
    }
    float inner(const PLSpectrum &s) const {  // :104-112, summed in FP32 over integer nm
        float sum = 0;
        for (int l = kLambdaMin; l <= kLambdaMax; ++l) sum += (*this)(float(l)) * s(float(l));
        return sum;
    }
};

// create_from_interleaved (:52-87): extend to [359, 831] if the data stops short.
inline PLSpectrum pl_from_interleaved(const float *s, int len) {
    PLSpectrum p;
    int n = len / 2;
    if (s[0] > kLambdaMin) {
        p.lam.push_back(kLambdaMin - 1);
        p.val.push_back(s[1]);
    }
    for (int i = 0; i < n; ++i) {
        p.lam.push_back(s[2 * i]);
        p.val.push_back(s[2 * i + 1]);
    }
    if (p.lam.back() < kLambdaMax) {
        p.lam.push_back(kLambdaMax + 1);
        p.val.push_back(p.val.back());
    }
    return p;
}

// 3x3 matrices (euclidean_space/squared_matrix.{h,cu}) with the reference's
// FMA-chained product and determinant-based inverse.
struct M3 {
    float m[3][3];
};
inline M3 m3_mul(const M3 &a, const M3 &b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            r.m[i][j] = 0;
            for (int k = 0; k < 3; ++k) r.m[i][j] = std::fma(a.m[i][k], b.m[k][j], r.m[i][j]);
        }
    return r;
}
inline M3 m3_inverse(const M3 &a) {
    const auto &m = a.m;
    float minor12 = dop(m[1][1], m[2][2], m[1][2], m[2][1]);
    float minor02 = dop(m[1][0], m[2][2], m[1][2], m[2][0]);
    float minor01 = dop(m[1][0], m[2][1], m[1][1], m[2][0]);
    float det = std::fma(m[0][2], minor01, dop(m[0][0], minor12, m[0][1], minor02));
    if (det == 0) throw std::runtime_error("singular 3x3 matrix");
    float inv_det = float(1.0 / double(det));
    M3 r;
    r.m[0][0] = inv_det * dop(m[1][1], m[2][2], m[1][2], m[2][1]);
    r.m[1][0] = inv_det * dop(m[1][2], m[2][0], m[1][0], m[2][2]);
    r.m[2][0] = inv_det * dop(m[1][0], m[2][1], m[1][1], m[2][0]);
    r.m[0][1] = inv_det * dop(m[0][2], m[2][1], m[0][1], m[2][2]);
    r.m[1][1] = inv_det * dop(m[0][0], m[2][2], m[0][2], m[2][0]);
    r.m[2][1] = inv_det * dop(m[0][1], m[2][0], m[0][0], m[2][1]);
    r.m[0][2] = inv_det * dop(m[0][1], m[1][2], m[0][2], m[1][1]);
    r.m[1][2] = inv_det * dop(m[0][2], m[1][0], m[0][0], m[1][2]);
    r.m[2][2] = inv_det * dop(m[0][0], m[1][1], m[0][1], m[1][0]);
    return r;
}

// util/compensated_float.h — Graillat's compensated dot product, evaluated from
// the last pair backwards as the reference's recursive template does.
inline float inner_product_n(const float *a, const float *b, int n) {
    float v = a[n - 1] * b[n - 1];
    float err = std::fma(a[n - 1], b[n - 1], -v);
    for (int i = n - 2; i >= 0; --i) {
        float ab = a[i] * b[i];
        float ab_err = std::fma(a[i], b[i], -ab);
        float s = ab + v;
        float delta = s - ab;
        float s_err = (ab - (s - delta)) + (v - delta);
        err = ab_err + (err + s_err);
        v = s;
    }
    return v + err;
}
inline std::array<float, 3> m3_apply(const M3 &m, const float v[3]) {  // rgb.h:99-103, xyz.h:102-106
    std::array<float, 3> r;
    for (int i = 0; i < 3; ++i) r[i] = inner_product_n(m.m[i], v, 3);
    return r;
}

struct Sigmoid {  // rgb_sigmoid_polynomial.h
    float c0 = 0, c1 = 0, c2 = 0;
    float operator()(float l) const {
        float x = std::fma(l, std::fma(l, c0, c1), c2);
        if (std::isinf(x)) return x > 0 ? 1.f : 0.f;
        return .5f + x / (2 * std::sqrt(1 + x * x));
    }
};

// ---------------------------------------------------------------- Jakob-Hanika
namespace rgb2spec {

constexpr int kRes = 64;
constexpr int kCieSamples = 95;
constexpr int kFine = (kCieSamples - 1) * 3 + 1;
constexpr double kLMin = 360, kLMax = 830;
constexpr double kEps = 1e-4;

const double xyz_to_srgb[3][3] = {
    {3.240479, -1.537150, -0.498535}, {-0.969256, 1.875991, 0.041556}, {0.055648, -0.204043, 1.057311}};
const double srgb_to_xyz[3][3] = {
    {0.412453, 0.357580, 0.180423}, {0.212671, 0.715160, 0.072169}, {0.019334, 0.119193, 0.950227}};

struct Buffer {
    double lambda[kFine];
    double rgb_tbl[3][kFine];
    double xyz_w[3];
};

inline double sigmoid(double x) { return 0.5 * x / std::sqrt(1.0 + x * x) + 0.5; }
inline double smoothstep(double x) { return x * x * (3.0 - 2.0 * x); }

inline double cie_interp(const double *data, double x) {
    x -= kLMin;
    x *= (kCieSamples - 1) / (kLMax - kLMin);
    int off = (int)x;
    if (off < 0) off = 0;
    if (off > kCieSamples - 2) off = kCieSamples - 2;
    double w = x - off;
    return (1.0 - w) * data[off] + w * data[off + 1];
}

inline void init_buffer(Buffer &b) {
    std::memset(b.rgb_tbl, 0, sizeof(b.rgb_tbl));
    std::memset(b.xyz_w, 0, sizeof(b.xyz_w));
    double h = (kLMax - kLMin) / double(kFine - 1);
    for (int i = 0; i < kFine; ++i) {
        double lambda = kLMin + i * h;
        double xyz[3] = {cie_interp(pbrtm_data::kCieX5nm, lambda), cie_interp(pbrtm_data::kCieY5nm, lambda),
                         cie_interp(pbrtm_data::kCieZ5nm, lambda)};
        double I = cie_interp(pbrtm_data::kD65_5nm, lambda);
        double ratio = (i == 0 || i == kFine - 1) ? 1.0 : (((i - 1) % 3 == 2) ? 2.0 : 3.0);
        double weight = 3.0 / 8.0 * h * ratio;
        b.lambda[i] = lambda;
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) b.rgb_tbl[k][i] += xyz_to_srgb[k][j] * xyz[j] * I * weight;
        for (int k = 0; k < 3; ++k) b.xyz_w[k] += xyz[k] * I * weight;
    }
}

inline void cie_lab(double *p, const Buffer &b) {
    double X = 0, Y = 0, Z = 0;
    for (int j = 0; j < 3; ++j) {
        X += p[j] * srgb_to_xyz[0][j];
        Y += p[j] * srgb_to_xyz[1][j];
        Z += p[j] * srgb_to_xyz[2][j];
    }
    auto f = [](double t) {
        double delta = 6.0 / 29.0;
        if (t > delta * delta * delta) return std::cbrt(t);
        return t / (delta * delta * 3.0) + (4.0 / 29.0);
    };
    p[0] = 116.0 * f(Y / b.xyz_w[1]) - 16.0;
    p[1] = 500.0 * (f(X / b.xyz_w[0]) - f(Y / b.xyz_w[1]));
    p[2] = 200.0 * (f(Y / b.xyz_w[1]) - f(Z / b.xyz_w[2]));
}

inline void residual(const Buffer &b, const double *c, const double *rgb, double *res) {
    double out[3] = {0, 0, 0};
    for (int i = 0; i < kFine; ++i) {
        double l = (b.lambda[i] - kLMin) / (kLMax - kLMin);
        double x = 0.0;
        for (int k = 0; k < 3; ++k) x = x * l + c[k];
        double s = sigmoid(x);
        for (int j = 0; j < 3; ++j) out[j] += b.rgb_tbl[j][i] * s;
    }
    cie_lab(out, b);
    std::memcpy(res, rgb, sizeof(double) * 3);
    cie_lab(res, b);
    for (int j = 0; j < 3; ++j) res[j] -= out[j];
}

// LU decomposition with partial pivoting by row-pointer swaps (rgb_to_spectrum_data.cu:60-136).
inline bool lup_solve(double **A, const double *rhs, double *x) {
    const int N = 3;
    int P[4] = {0, 1, 2, 3};
    for (int i = 0; i < N; ++i) {
        double maxA = 0.0;
        int imax = i;
        for (int k = i; k < N; ++k) {
            double a = std::fabs(A[k][i]);
            if (a > maxA) {
                maxA = a;
                imax = k;
            }
        }
        if (maxA < 1e-15) return false;
        if (imax != i) {
            std::swap(P[i], P[imax]);
            std::swap(A[i], A[imax]);
            P[N]++;
        }
        for (int j = i + 1; j < N; ++j) {
            A[j][i] /= A[i][i];
            for (int k = i + 1; k < N; ++k) A[j][k] -= A[j][i] * A[i][k];
        }
    }
    for (int i = 0; i < N; ++i) {
        x[i] = rhs[P[i]];
        for (int k = 0; k < i; ++k) x[i] -= A[i][k] * x[k];
    }
    for (int i = N - 1; i >= 0; --i) {
        for (int k = i + 1; k < N; ++k) x[i] -= A[i][k] * x[k];
        x[i] = x[i] / A[i][i];
    }
    return true;
}

inline void gauss_newton(const Buffer &b, const double rgb[3], double c[3]) {
    for (int it = 0; it < 15; ++it) {
        double J0[3], J1[3], J2[3], *J[3] = {J0, J1, J2};
        double r[3];
        residual(b, c, rgb, r);
        double r0[3], r1[3], tmp[3];
        for (int i = 0; i < 3; ++i) {
            std::memcpy(tmp, c, sizeof(tmp));
            tmp[i] -= kEps;
            residual(b, tmp, rgb, r0);
            std::memcpy(tmp, c, sizeof(tmp));
            tmp[i] += kEps;
            residual(b, tmp, rgb, r1);
            for (int j = 0; j < 3; ++j) J[j][i] = (r1[j] - r0[j]) * 1.0 / (2 * kEps);
        }
        double x[3];
        if (!lup_solve(J, r, x)) throw std::runtime_error("LU decomposition failed");
        double rr = 0.0;
        for (int j = 0; j < 3; ++j) {
            c[j] -= x[j];
            rr += r[j] * r[j];
        }
        double mx = std::max(std::max(c[0], c[1]), c[2]);
        if (mx > 200)
            for (int j = 0; j < 3; ++j) c[j] *= 200 / mx;
        if (rr < 1e-6) break;
    }
}

// One column (l, j, i) of the coefficient table for every z index k, by the
// warm-started chains of compute() (rgb_to_spectrum_data.cu:305-358).
inline void column(const Buffer &b, const double *z_nodes, int l, int j, int i, double out[kRes][3]) {
    const double y = j / double(kRes - 1), x = i / double(kRes - 1);
    const double c0 = kLMin, c1 = 1.0 / (kLMax - kLMin);
    const int start = kRes / 5;
    auto store = [&](int k, const double *c) {
        double A = c[0], B = c[1], C = c[2];
        out[k][0] = A * (c1 * c1);
        out[k][1] = B * c1 - 2 * A * c0 * (c1 * c1);
        out[k][2] = C - B * c0 * c1 + A * ((c0 * c1) * (c0 * c1));
    };
    double coeffs[3] = {0, 0, 0}, rgb[3];
    for (int k = start; k < kRes; ++k) {
        double bb = z_nodes[k];
        rgb[l] = bb;
        rgb[(l + 1) % 3] = x * bb;
        rgb[(l + 2) % 3] = y * bb;
        gauss_newton(b, rgb, coeffs);
        store(k, coeffs);
    }
    std::memset(coeffs, 0, sizeof(coeffs));
    for (int k = start; k >= 0; --k) {
        double bb = z_nodes[k];
        rgb[l] = bb;
        rgb[(l + 1) % 3] = x * bb;
        rgb[(l + 2) % 3] = y * bb;
        gauss_newton(b, rgb, coeffs);
        store(k, coeffs);
    }
}

class Table {
  public:
    Table() {
        init_buffer(buf_);
        for (int k = 0; k < kRes; ++k) z_nodes_[k] = smoothstep(smoothstep(double(k) / double(kRes - 1)));
    }
    const std::array<std::array<double, 3>, kRes> &col(int l, int j, int i) {
        std::lock_guard<std::mutex> g(mu_);
        int key = (l * kRes + j) * kRes + i;
        auto it = cols_.find(key);
        if (it != cols_.end()) return it->second;
        double tmp[kRes][3];
        column(buf_, z_nodes_, l, j, i, tmp);
        auto &dst = cols_[key];
        for (int k = 0; k < kRes; ++k)
            for (int q = 0; q < 3; ++q) dst[k][q] = tmp[k][q];
        return dst;
    }
    // RGBtoSpectrumTable::operator() (rgb_to_spectrum_data.cu:361-400).
    Sigmoid lookup(const float rgb[3]) {
        if (rgb[0] == rgb[1] && rgb[1] == rgb[2]) {
            Sigmoid s;
            s.c2 = float(double(rgb[0] - 0.5f) / std::sqrt(double(rgb[0]) * (1.0 - double(rgb[0]))));
            return s;
        }
        int mc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0 : 2) : ((rgb[1] > rgb[2]) ? 1 : 2);
        double z = rgb[mc];
        double x = double(rgb[(mc + 1) % 3] * (kRes - 1)) / z;
        double y = double(rgb[(mc + 2) % 3] * (kRes - 1)) / z;
        int xi = std::min((int)x, kRes - 2), yi = std::min((int)y, kRes - 2);
        int zi = (int)find_interval(kRes, [&](size_t i) { return z_nodes_[i] < z; });
        float dx = float(x - xi), dy = float(y - yi);
        float dz = float((z - z_nodes_[zi]) / (z_nodes_[zi + 1] - z_nodes_[zi]));
        float c[3];
        const auto &c00 = col(mc, yi, xi), &c10 = col(mc, yi, xi + 1);
        const auto &c01 = col(mc, yi + 1, xi), &c11 = col(mc, yi + 1, xi + 1);
        for (int q = 0; q < 3; ++q) {
            auto co = [&](int ddx, int ddy, int ddz) -> float {
                const auto &cc = ddy ? (ddx ? c11 : c01) : (ddx ? c10 : c00);
                return float(cc[zi + ddz][q]);
            };
            c[q] = lerpf(dz, lerpf(dy, lerpf(dx, co(0, 0, 0), co(1, 0, 0)), lerpf(dx, co(0, 1, 0), co(1, 1, 0))),
                         lerpf(dy, lerpf(dx, co(0, 0, 1), co(1, 0, 1)), lerpf(dx, co(0, 1, 1), co(1, 1, 1))));
        }
        return Sigmoid{c[0], c[1], c[2]};
    }

    // d(c0, c1, c2)/d(r, g, b) of lookup(): J[q][k] = dc_q / drgb_k, from the
    // derivative of the trilinear interpolation and of the cell coordinates
    // x = rgb[(m+1)%3] (res-1) / z, y = rgb[(m+2)%3] (res-1) / z, z = rgb[m].
    // Grey inputs take the general branch (the closed form holds only on the diagonal).
    void jacobian(const float rgb[3], double J[3][3]) {
        int mc = (rgb[0] > rgb[1]) ? ((rgb[0] > rgb[2]) ? 0 : 2) : ((rgb[1] > rgb[2]) ? 1 : 2);
        double z = rgb[mc];
        double x = double(rgb[(mc + 1) % 3] * (kRes - 1)) / z;
        double y = double(rgb[(mc + 2) % 3] * (kRes - 1)) / z;
        int xi = std::min((int)x, kRes - 2), yi = std::min((int)y, kRes - 2);
        int zi = (int)find_interval(kRes, [&](size_t i) { return z_nodes_[i] < z; });
        double dx = x - xi, dy = y - yi;
        double dzn = z_nodes_[zi + 1] - z_nodes_[zi];
        double dz = (z - z_nodes_[zi]) / dzn;
        const auto &c00 = col(mc, yi, xi), &c10 = col(mc, yi, xi + 1);
        const auto &c01 = col(mc, yi + 1, xi), &c11 = col(mc, yi + 1, xi + 1);
        auto lerp = [](double t, double a, double b) { return (1 - t) * a + t * b; };
        for (int q = 0; q < 3; ++q) {
            auto co = [&](int ddx, int ddy, int ddz) -> double {
                const auto &cc = ddy ? (ddx ? c11 : c01) : (ddx ? c10 : c00);
                return double(float(cc[zi + ddz][q]));
            };
            const double gx = lerp(dz, lerp(dy, co(1, 0, 0) - co(0, 0, 0), co(1, 1, 0) - co(0, 1, 0)),
                                   lerp(dy, co(1, 0, 1) - co(0, 0, 1), co(1, 1, 1) - co(0, 1, 1)));
            const double gy = lerp(dz, lerp(dx, co(0, 1, 0) - co(0, 0, 0), co(1, 1, 0) - co(1, 0, 0)),
                                   lerp(dx, co(0, 1, 1) - co(0, 0, 1), co(1, 1, 1) - co(1, 0, 1)));
            const double gz = lerp(dy, lerp(dx, co(0, 0, 1), co(1, 0, 1)), lerp(dx, co(0, 1, 1), co(1, 1, 1))) -
                              lerp(dy, lerp(dx, co(0, 0, 0), co(1, 0, 0)), lerp(dx, co(0, 1, 0), co(1, 1, 0)));
            J[q][mc] = gx * (-x / z) + gy * (-y / z) + gz / dzn;
            J[q][(mc + 1) % 3] = gx * (kRes - 1) / z;
            J[q][(mc + 2) % 3] = gy * (kRes - 1) / z;
        }
    }

  private:
    Buffer buf_;
    double z_nodes_[kRes];
    std::map<int, std::array<std::array<double, 3>, kRes>> cols_;
    std::mutex mu_;
};

}  // namespace rgb2spec

// GlobalSpectra + sRGB colour space + RGB film sensor constants.
struct Spectra {
    PLSpectrum cie[3];  // with the reference's 471 leading (0, 0) pairs (global_spectra.cu:10-21)
    PLSpectrum d65;     // normalised to CIE_Y_integral
    M3 xyz_from_rgb, rgb_from_xyz;
    float white_xy[2];
    rgb2spec::Table table;

    Spectra() {
        const float *cie_v[3] = {pbrtm_data::kCieX1nm, pbrtm_data::kCieY1nm, pbrtm_data::kCieZ1nm};
        for (int c = 0; c < 3; ++c) {
            cie[c].lam.assign(471, 0.f);
            cie[c].val.assign(471, 0.f);
            for (int i = 0; i < 471; ++i) {
                cie[c].lam.push_back(float(kLambdaMin + i));
                cie[c].val.push_back(cie_v[c][i]);
            }
        }
        d65 = pl_from_interleaved(pbrtm_data::kD65Interleaved, pbrtm_data::kD65InterleavedLen);
        const float s = pbrtm_data::kCieYIntegral / d65.inner(cie[1]);
        for (auto &v : d65.val) v *= s;
        // RGBColorSpace::init with r (0.64, 0.33), g (0.3, 0.6), b (0.15, 0.06).
        float w[3];
        for (int c = 0; c < 3; ++c) w[c] = d65.inner(cie[c]) / pbrtm_data::kCieYIntegral;
        white_xy[0] = w[0] / (w[0] + w[1] + w[2]);
        white_xy[1] = w[1] / (w[0] + w[1] + w[2]);
        const float prim[3][2] = {{0.64f, 0.33f}, {0.3f, 0.6f}, {0.15f, 0.06f}};
        M3 rgb;
        for (int c = 0; c < 3; ++c) {
            float x = prim[c][0], y = prim[c][1];
            rgb.m[0][c] = x * 1.0f / y;
            rgb.m[1][c] = 1.0f;
            rgb.m[2][c] = (1 - x - y) * 1.0f / y;
        }
        auto cw = m3_apply(m3_inverse(rgb), w);
        M3 diag{};
        for (int c = 0; c < 3; ++c) diag.m[c][c] = cw[c];
        xyz_from_rgb = m3_mul(rgb, diag);
        rgb_from_xyz = m3_inverse(xyz_from_rgb);
    }
};

}  // namespace orc
