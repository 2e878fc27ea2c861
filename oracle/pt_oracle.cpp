// TEST INFRASTRUCTURE ONLY — part of the CPU oracle (see oracle_pt.h).
//
// Host-only restatement of the reference's wavefront path tracer for diffuse
// and conductor surfaces lit by diffuse area lights. Every FP32 expression keeps
// the reference's operation order and its double-precision promotions; the
// build uses -ffp-contract=off so a*b+c is two roundings unless std::fma is
// written out. Transcendental functions (sin, cos, asin, atan2, atanh, cosh,
// tan) are evaluated in double precision and rounded once to float, a
// correctly rounded stand-in for the CUDA device library the reference calls.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_math.h"
#include "oracle_pt.h"
#include "oracle_quadric.h"
#include "oracle_rng.h"
#include "oracle_spectra.h"
#include "oracle_tri.h"

namespace orc {
namespace {

// ------------------------------------------------------------ small helpers
inline float tsin(float x) { return float(std::sin(double(x))); }
inline float tcos(float x) { return float(std::cos(double(x))); }
inline float tasin(float x) { return float(std::asin(double(x))); }
inline float tatan2(float y, float x) { return float(std::atan2(double(y), double(x))); }
inline float tatanh(float x) { return float(std::atanh(double(x))); }
inline float tcosh(float x) { return float(std::cosh(double(x))); }

const float kPi = float(std::acos(-1.0));  // compute_pi(): acos(-1) rounded to float
constexpr float kOneMinusEps = 0x1.fffffep-1f;

inline float sqrf(float v) { return v * v; }
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float safe_sqrt(float x) { return std::sqrt(std::max(0.0f, x)); }
inline float safe_asin(float x) { return tasin(clampf(x, -1, 1)); }
inline float length(V3 a) { return std::sqrt(sqlen(a)); }
inline V3 vmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float abs_dot(V3 a, V3 b) { return std::fabs(dot(a, b)); }
inline bool v3_zero(V3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }
inline V3 face_forward(V3 n, V3 v) { return dot(n, v) < 0.0f ? neg(n) : n; }

// vector3.h:233-240
inline float angle_between(V3 v1, V3 v2) {
    if (dot(v1, v2) < 0) return kPi - 2 * safe_asin(length(add(v1, v2)) / 2);
    return 2 * safe_asin(length(sub(v2, v1)) / 2);
}
inline V3 gram_schmidt(V3 v, V3 w) { return sub(v, scale(dot(v, w), w)); }

// ------------------------------------------------------------ spectra (4 wavelengths)
struct S4 {
    float v[kNS];
};
inline S4 s4(float a) { return {{a, a, a, a}}; }
inline S4 operator*(const S4 &a, const S4 &b) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = a.v[i] * b.v[i];
    return r;
}
inline S4 operator*(const S4 &a, float s) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = a.v[i] * s;
    return r;
}
inline S4 operator/(const S4 &a, float s) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = a.v[i] / s;
    return r;
}
inline S4 operator+(const S4 &a, const S4 &b) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = a.v[i] + b.v[i];
    return r;
}
inline bool positive(const S4 &a) {
    for (int i = 0; i < kNS; ++i)
        if (a.v[i] > 0) return true;
    return false;
}
inline float max_comp(const S4 &a) {
    float m = a.v[0];
    for (int i = 1; i < kNS; ++i) m = std::max(m, a.v[i]);
    return m;
}
inline float average(const S4 &a) {
    float s = 0;
    for (int i = 0; i < kNS; ++i) s += a.v[i];
    return s / kNS;
}
inline S4 safe_div(const S4 &a, const S4 &b) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = b.v[i] == 0.0f ? 0.0f : a.v[i] / b.v[i];
    return r;
}

struct Lambda {
    float l[kNS], pdf[kNS];
};
// sampling.h:240-253, sampled_wavelengths.h:25-42
inline Lambda sample_visible(float u) {
    Lambda w;
    for (int i = 0; i < kNS; ++i) {
        float up = u + float(i) / kNS;
        if (up > 1) up -= 1;
        float l = 538 - 138.888889f * tatanh(0.85691062f - 1.82750197f * up);
        w.l[i] = l;
        w.pdf[i] = (l < kLambdaMin || l > kLambdaMax) ? 0.0f : 0.0039398042f / sqrf(tcosh(0.0072f * (l - 538)));
    }
    return w;
}
inline S4 sample_pl(const PLSpectrum &s, const Lambda &w) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = s(w.l[i]);
    return r;
}
inline S4 sample_sigmoid(const Sigmoid &s, const Lambda &w) {
    S4 r;
    for (int i = 0; i < kNS; ++i) r.v[i] = s(w.l[i]);
    return r;
}

// ------------------------------------------------------------ 4x4 transforms
struct M4 {
    float m[4][4];
};
inline M4 m4_mul(const M4 &a, const M4 &b) {  // squared_matrix.h:98-109
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            r.m[i][j] = 0;
            for (int k = 0; k < 4; ++k) r.m[i][j] = std::fma(a.m[i][k], b.m[k][j], r.m[i][j]);
        }
    return r;
}
inline float ip(std::initializer_list<float> t) {
    std::vector<float> a, b;
    auto it = t.begin();
    while (it != t.end()) {
        a.push_back(*it++);
        b.push_back(*it++);
    }
    return inner_product_n(a.data(), b.data(), int(a.size()));
}
inline M4 m4_inverse(const M4 &A) {  // squared_matrix.cu (Eberly's Laplace expansion)
    const auto &m = A.m;
    float s0 = dop(m[0][0], m[1][1], m[1][0], m[0][1]);
    float s1 = dop(m[0][0], m[1][2], m[1][0], m[0][2]);
    float s2 = dop(m[0][0], m[1][3], m[1][0], m[0][3]);
    float s3 = dop(m[0][1], m[1][2], m[1][1], m[0][2]);
    float s4_ = dop(m[0][1], m[1][3], m[1][1], m[0][3]);
    float s5 = dop(m[0][2], m[1][3], m[1][2], m[0][3]);
    float c0 = dop(m[2][0], m[3][1], m[3][0], m[2][1]);
    float c1 = dop(m[2][0], m[3][2], m[3][0], m[2][2]);
    float c2 = dop(m[2][0], m[3][3], m[3][0], m[2][3]);
    float c3 = dop(m[2][1], m[3][2], m[3][1], m[2][2]);
    float c4 = dop(m[2][1], m[3][3], m[3][1], m[2][3]);
    float c5 = dop(m[2][2], m[3][3], m[3][2], m[2][3]);
    float det = ip({s0, c5, -s1, c4, s2, c3, s3, c2, s5, c0, -s4_, c1});
    if (det == 0) throw std::runtime_error("singular 4x4 matrix");
    float s = float(1.0 / double(det));
    M4 r;
    r.m[0][0] = s * ip({m[1][1], c5, m[1][3], c3, -m[1][2], c4});
    r.m[0][1] = s * ip({-m[0][1], c5, m[0][2], c4, -m[0][3], c3});
    r.m[0][2] = s * ip({m[3][1], s5, m[3][3], s3, -m[3][2], s4_});
    r.m[0][3] = s * ip({-m[2][1], s5, m[2][2], s4_, -m[2][3], s3});
    r.m[1][0] = s * ip({-m[1][0], c5, m[1][2], c2, -m[1][3], c1});
    r.m[1][1] = s * ip({m[0][0], c5, m[0][3], c1, -m[0][2], c2});
    r.m[1][2] = s * ip({-m[3][0], s5, m[3][2], s2, -m[3][3], s1});
    r.m[1][3] = s * ip({m[2][0], s5, m[2][3], s1, -m[2][2], s2});
    r.m[2][0] = s * ip({m[1][0], c4, m[1][3], c0, -m[1][1], c2});
    r.m[2][1] = s * ip({-m[0][0], c4, m[0][1], c2, -m[0][3], c0});
    r.m[2][2] = s * ip({m[3][0], s4_, m[3][3], s0, -m[3][1], s2});
    r.m[2][3] = s * ip({-m[2][0], s4_, m[2][1], s2, -m[2][3], s0});
    r.m[3][0] = s * ip({-m[1][0], c3, m[1][1], c1, -m[1][2], c0});
    r.m[3][1] = s * ip({m[0][0], c3, m[0][2], c0, -m[0][1], c1});
    r.m[3][2] = s * ip({-m[3][0], s3, m[3][1], s1, -m[3][2], s0});
    r.m[3][3] = s * ip({m[2][0], s3, m[2][2], s0, -m[2][1], s1});
    return r;
}
struct Xf {  // euclidean_space/transform.h
    M4 m, inv;
};
inline M4 m4_diag(float a, float b, float c) {
    M4 r{};
    r.m[0][0] = a;
    r.m[1][1] = b;
    r.m[2][2] = c;
    r.m[3][3] = 1;
    return r;
}
inline Xf xf_from(const M4 &m) { return {m, m4_inverse(m)}; }
inline Xf xf_mul(const Xf &a, const Xf &b) { return {m4_mul(a.m, b.m), m4_mul(b.inv, a.inv)}; }
inline Xf xf_inverse(const Xf &a) { return {a.inv, a.m}; }
inline Xf xf_scale(float x, float y, float z) {
    return {m4_diag(x, y, z), m4_diag(float(1.0) / x, float(1.0) / y, float(1.0) / z)};
}
inline Xf xf_translate(float x, float y, float z) {
    Xf t{m4_diag(1, 1, 1), m4_diag(1, 1, 1)};
    t.m.m[0][3] = x;
    t.m.m[1][3] = y;
    t.m.m[2][3] = z;
    t.inv.m[0][3] = -x;
    t.inv.m[1][3] = -y;
    t.inv.m[2][3] = -z;
    return t;
}
inline V3 xf_point(const M4 &M, V3 p) {  // transform.h:254-265
    const auto &m = M.m;
    float xp = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    float yp = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    float zp = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    float wp = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (wp == 1.0f) return {xp, yp, zp};
    return {xp / wp, yp / wp, zp / wp};
}
inline V3 xf_vector(const M4 &M, V3 v) {  // transform.h:303-307
    const auto &m = M.m;
    return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}

// transform.h:181-224 — camera_from_world with world_from_camera as its inverse.
inline Xf look_at(V3 pos, V3 look, V3 up) {
    M4 w{};
    w.m[0][3] = pos.x;
    w.m[1][3] = pos.y;
    w.m[2][3] = pos.z;
    w.m[3][3] = 1.0f;
    V3 dir = normalize(sub(look, pos));
    if (length(cross(normalize(up), dir)) == 0.0f) throw std::runtime_error("LookAt: up parallel to view");
    V3 right = normalize(cross(normalize(up), dir));
    V3 new_up = cross(dir, right);
    w.m[0][0] = right.x;
    w.m[1][0] = right.y;
    w.m[2][0] = right.z;
    w.m[0][1] = new_up.x;
    w.m[1][1] = new_up.y;
    w.m[2][1] = new_up.z;
    w.m[0][2] = dir.x;
    w.m[1][2] = dir.y;
    w.m[2][2] = dir.z;
    return {m4_inverse(w), w};
}

// ------------------------------------------------------------ camera
struct Camera {
    Xf render_from_camera, camera_from_raster;
    Xf render_from_world;
};
// cameras/perspective.cu:6-61, base/camera.h CameraTransform (render space = camera-world).
inline Camera make_camera(const orc_scene &s) {
    Camera c;
    Xf camera_from_world = look_at({s.eye[0], s.eye[1], s.eye[2]}, {s.look[0], s.look[1], s.look[2]},
                                   {s.up[0], s.up[1], s.up[2]});
    Xf world_from_camera = xf_inverse(camera_from_world);
    V3 pc = xf_point(world_from_camera.m, {0, 0, 0});
    Xf world_from_render = xf_translate(pc.x, pc.y, pc.z);
    c.render_from_world = xf_inverse(world_from_render);
    c.render_from_camera = xf_mul(c.render_from_world, world_from_camera);
    float aspect = float(s.width) / float(s.height);
    float sw[4];  // pmin.x, pmin.y, pmax.x, pmax.y
    if (aspect > 1.0) {
        sw[0] = -aspect, sw[1] = -1.0f, sw[2] = aspect, sw[3] = 1.0f;
    } else {
        sw[0] = -1.0f, sw[1] = float(-1.0 / aspect), sw[2] = 1.0f, sw[3] = float(1.0 / aspect);
    }
    Xf ndc_from_screen = xf_mul(xf_scale(float(1.0 / (sw[2] - sw[0])), float(1.0 / (sw[3] - sw[1])), 1.0f),
                                xf_translate(-sw[0], -sw[3], 0.0f));
    Xf raster_from_ndc = xf_scale(float(s.width), float(-int(s.height)), 1.0f);
    Xf raster_from_screen = xf_mul(raster_from_ndc, ndc_from_screen);
    Xf screen_from_raster = xf_inverse(raster_from_screen);
    const float z_near = 1e-2f, z_far = 1000.0f;
    M4 persp{};
    persp.m[0][0] = 1;
    persp.m[1][1] = 1;
    persp.m[2][2] = z_far / (z_far - z_near);
    persp.m[2][3] = -z_far * z_near / (z_far - z_near);
    persp.m[3][2] = 1;
    float deg = float(double(kPi) / 180.0 * double(s.fov));
    double inv_tan = 1.0 / std::tan(double(deg) / 2.0);
    Xf screen_from_camera = xf_mul(xf_scale(float(inv_tan), float(inv_tan), 1.0f), xf_from(persp));
    c.camera_from_raster = xf_mul(xf_inverse(screen_from_camera), screen_from_raster);
    return c;
}

// ------------------------------------------------------------ Mitchell filter sampler
struct Pc1d {  // util/piecewise_constant_1d.h
    std::vector<float> func, cdf;
    float mn, mx, func_int;
    void init(const float *f, size_t n, float lo, float hi) {
        mn = lo;
        mx = hi;
        func.resize(n);
        for (size_t i = 0; i < n; ++i) func[i] = std::fabs(f[i]);
        cdf.assign(n + 1, 0.f);
        for (size_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] * (mx - mn) / float(n);
        func_int = cdf[n];
        if (func_int == 0)
            for (size_t i = 1; i < n + 1; ++i) cdf[i] = float(i) / float(n);
        else
            for (size_t i = 1; i < n + 1; ++i) cdf[i] /= func_int;
    }
    float sample(float u, float *pdf, int *off) const {
        int o = (int)find_interval(func.size() + 1, [&](size_t i) { return cdf[i] <= u; });
        *off = o;
        float du = u - cdf[o];
        if (cdf[o + 1] - cdf[o] > 0) du /= cdf[o + 1] - cdf[o];
        *pdf = (func_int > 0) ? func[o] / func_int : 0;
        return lerpf((float(o) + du) / float(func.size()), mn, mx);
    }
};
struct FilterSampler {  // base/filter.cu:146-161, filters/mitchell.{h,cu}
    float radius = 2.f, b = 1.f / 3.f, c = 1.f / 3.f;
    int nx = 64, ny = 64;
    std::vector<float> f;
    std::vector<Pc1d> cond;
    Pc1d marg;
    float mitchell_1d(float x) const {
        x = std::fabs(x);
        if (x <= 1)
            return ((12 - 9 * b - 6 * c) * x * x * x + (-18 + 12 * b + 6 * c) * x * x + (6 - 2 * b)) * (1.f / 6.f);
        if (x <= 2)
            return ((-b - 6 * c) * x * x * x + (6 * b + 30 * c) * x * x + (-12 * b - 48 * c) * x + (8 * b + 24 * c)) *
                   (1.f / 6.f);
        return 0;
    }
    FilterSampler() {
        nx = int(32 * radius);
        ny = int(32 * radius);
        f.resize(size_t(nx) * ny);
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float px = lerpf((x + 0.5f) / nx, -radius, radius);
                float py = lerpf((y + 0.5f) / ny, -radius, radius);
                f[x + nx * y] = mitchell_1d(2 * px / radius) * mitchell_1d(2 * py / radius);
            }
        cond.resize(ny);
        std::vector<float> mf;
        for (int v = 0; v < ny; ++v) {
            cond[v].init(&f[size_t(v) * nx], nx, -radius, radius);
            mf.push_back(cond[v].func_int);
        }
        marg.init(mf.data(), mf.size(), -radius, radius);
    }
    void sample(float u0, float u1, float *px, float *py, float *w) const {
        float pdf0, pdf1;
        int o0, o1;
        float d1 = marg.sample(u1, &pdf1, &o1);
        float d0 = cond[o1].sample(u0, &pdf0, &o0);
        *px = d0;
        *py = d1;
        *w = f[o0 + nx * o1] / (pdf0 * pdf1);
    }
};

// ------------------------------------------------------------ stratified sampler
inline uint64_t hash_u32(uint32_t a) { return murmur64a(reinterpret_cast<const unsigned char *>(&a), 4, 0); }
inline uint64_t hash_u32x2(uint32_t a, uint32_t b) {
    uint32_t buf[2] = {a, b};
    return murmur64a(reinterpret_cast<const unsigned char *>(buf), 8, 0);
}
inline int permutation_element(uint32_t i, uint32_t l, uint32_t p) {  // util/permutation.h (Kensler 2013)
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
    return int((i + p) % l);
}
struct Sampler {  // samplers/stratified.h
    uint32_t spd, pixel, sample, dim;
    Pcg32 rng;
    void start(uint32_t px, uint32_t s, uint32_t d) {
        pixel = px;
        sample = s;
        dim = d;
        rng.set_sequence(hash_u32(px));
        rng.advance(int64_t(uint64_t(s) * 65536ull + d));
    }
    float get_1d() {
        uint64_t h = hash_u32x2(pixel, dim);
        int stratum = permutation_element(sample, spd * spd, uint32_t(h));
        dim += 1;
        float delta = rng.next_float();
        return (float(stratum) + delta) / float(spd * spd);
    }
    void get_2d(float *u0, float *u1) {
        uint64_t h = hash_u32x2(pixel, dim);
        int stratum = permutation_element(sample, spd * spd, uint32_t(h));
        dim += 2;
        uint32_t x = uint32_t(stratum) % spd, y = uint32_t(stratum) / spd;
        float dx = rng.next_float();
        float dy = rng.next_float();
        *u0 = (float(x) + dx) / float(spd);
        *u1 = (float(y) + dy) / float(spd);
    }
};

// ------------------------------------------------------------ scene tables
struct Tri {
    V3 p[3];
    uint32_t material;
    int32_t light;
    bool flip;
};
struct Material {
    uint32_t type;
    Sigmoid albedo;               // diffuse
    const PLSpectrum *eta, *k;    // conductor
    float eta_c;                  // dielectric, constant eta
    const PLSpectrum *eta_pl;     // dielectric, spectral eta (secondary wavelengths terminated)
    float ax, ay;
};
struct Light {
    uint32_t tri;
    float scale;                  // light scale / to_photometric
    float rgb_scale;              // RGBIlluminantSpectrum::scale
    Sigmoid rsp;
    bool two_sided;
    bool infinite = false;        // UniformInfiniteLight (tri unused; no "L": rgb_scale 1, rsp = 1)
};

struct Scene {
    const orc_scene *desc;
    Camera cam;
    FilterSampler filter;
    std::vector<Tri> tris;
    std::vector<float> soup;      // render space, for the oracle BVH
    orc_bvh *bvh = nullptr;
    std::vector<Material> mats;
    std::vector<Light> lights;
    std::vector<float> light_pmf, light_cdf;
    std::vector<float> select_cdf;  // the cdf the NEE light choice searches (= light_cdf unless overridden)
    std::vector<uint32_t> infinite;  // indices of the infinite lights, in light order
    float inf_dist = 0;             // 2 x the scene's bounding-sphere radius
    std::vector<quad::Quadric> quads;  // analytic primitives (spheres, disks, instanced triangles)
    std::vector<uint32_t> prim_quad;   // per primitive: quads index + 1, 0 = triangle (empty: none)
    std::vector<uint8_t> is_quad;      // the oracle BVH's custom-test flags
    Spectra *sp;
    PLSpectrum metals[8], glass_bk7;
    M3 out_from_sensor;
    ~Scene() {
        if (bvh) orc_bvh_free(bvh);
    }
};

Spectra &spectra() {
    static Spectra s;
    return s;
}

inline float tri_area(const Tri &t) { return length(cross(sub(t.p[1], t.p[0]), sub(t.p[2], t.p[0]))) * 0.5f; }

// lights/diffuse_area_light.cu:37-46
inline S4 light_L(const Scene &sc, const Light &L, V3 n, V3 w, const Lambda &lam) {
    if (!L.two_sided && dot(n, w) < 0.0f) return s4(0.0f);
    S4 s;
    for (int i = 0; i < kNS; ++i) s.v[i] = L.rgb_scale * L.rsp(lam.l[i]);
    return (sample_pl(sc.sp->d65, lam) * s) * L.scale;
}

// Transform::is_identity (transform.h:219-239)
inline bool xf_is_identity(const Xf &x) {
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            if (x.m.m[i][k] != x.inv.m[i][k] || x.m.m[i][k] != (i == k ? 1.f : 0.f)) return false;
    return true;
}
// Transform::operator()(Bounds3f) (transform.h:323-349): translation, then per
// output axis the smaller / larger of the two scaled input planes.
inline B3 xf_bounds(const M4 &M, const B3 &b) {
    B3 r;
    for (int i = 0; i < 3; ++i) r.lo[i] = r.hi[i] = M.m[i][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            const float a = M.m[i][k] * b.lo[k], c = M.m[i][k] * b.hi[k];
            r.lo[i] += std::min(a, c);
            r.hi[i] += std::max(a, c);
        }
    return r;
}
// SquareMatrix<3>::determinant (squared_matrix.cu:8-17) of the upper 3x3 < 0.
inline bool xf_swaps_handedness(const M4 &M) {
    const auto &m = M.m;
    const float minor12 = dop(m[1][1], m[2][2], m[1][2], m[2][1]);
    const float minor02 = dop(m[1][0], m[2][2], m[1][2], m[2][0]);
    const float minor01 = dop(m[1][0], m[2][1], m[1][1], m[2][0]);
    return std::fma(m[0][2], minor01, dop(m[0][0], minor12, m[0][1], minor02)) < 0;
}

// Entry i is an analytic shape or an instance's triangle (Shape::create,
// shapes/sphere.cu:8-40, shapes/disk.cu:6-25; ObjectInstance, scene_builder.cu:775-810).
// Its oracle-BVH entry is its bounds as the triangle (lo, hi, lo).
void add_quadric(Scene &sc, const orc_scene &d, uint64_t i, const Xf &render_from_world) {
    const orc_shape &sh = d.shapes[d.tri_shape[i]];
    if (d.tri_shape[i] >= d.num_shapes) throw std::runtime_error("shape index out of range");
    Xf ctm{m4_diag(1, 1, 1), m4_diag(1, 1, 1)};
    if (sh.xform != 0xFFFFFFFFu) {
        if (sh.xform >= d.num_xforms) throw std::runtime_error("shape transform index out of range");
        std::memcpy(ctm.m.m, d.xforms + 16 * size_t(sh.xform), sizeof(ctm.m.m));
        if (d.xforms_inv) std::memcpy(ctm.inv.m, d.xforms_inv + 16 * size_t(sh.xform), sizeof(ctm.inv.m));
        else ctm.inv = m4_inverse(ctm.m);
    }
    Xf x;
    Tri &t = sc.tris[i];
    quad::Quadric q{};
    q.type = sh.type;
    q.reverse = t.flip;
    if (sh.type == 1 || sh.type == 2) {
        x = xf_mul(render_from_world, ctm);
    } else if (sh.type == 3) {
        if (t.light >= 0) throw std::runtime_error("area lights with object instancing");
        x = xf_mul(xf_mul(render_from_world, ctm), xf_inverse(render_from_world));
        if (xf_is_identity(x)) return;  // the definition's triangle itself
    } else {
        throw std::runtime_error("unknown shape type");
    }
    for (const M4 *mm : {&x.m, &x.inv})
        if (mm->m[3][0] != 0 || mm->m[3][1] != 0 || mm->m[3][2] != 0 || mm->m[3][3] != 1)
            throw std::runtime_error("projective shape transform");
    std::memcpy(q.rfo.m, x.m.m, sizeof(q.rfo.m));
    std::memcpy(q.rfo.inv, x.inv.m, sizeof(q.rfo.inv));
    const float phi_max = float(double(kPi) / 180.0 * double(clampf(sh.phimax, 0, 360)));
    B3 b;
    if (sh.type == 1) {
        const float r = sh.radius;
        q.radius = r;
        q.z_min = clampf(std::min(sh.zmin, sh.zmax), -r, r);
        q.z_max = clampf(std::max(sh.zmin, sh.zmax), -r, r);
        q.theta_z_min = float(std::acos(double(clampf(q.z_min / r, -1, 1))));
        q.theta_z_max = float(std::acos(double(clampf(q.z_max / r, -1, 1))));
        q.phi_max = phi_max;
        q.flip = q.reverse;  // transform_swaps_handedness is never set by Sphere::init
        b = xf_bounds(x.m, B3{{-r, -r, q.z_min}, {r, r, q.z_max}});
    } else if (sh.type == 2) {
        q.height = sh.height;
        q.radius = sh.radius;
        q.inner_radius = sh.innerradius;
        q.phi_max = phi_max;
        q.flip = q.reverse ^ xf_swaps_handedness(x.m);
        b = xf_bounds(x.m, B3{{-sh.radius, -sh.radius, sh.height}, {sh.radius, sh.radius, sh.height}});
    } else {
        for (int v = 0; v < 3; ++v) q.p[v] = t.p[v];
        q.flip = q.reverse;
        b = xf_bounds(x.m, b3_of_tri(t.p[0], t.p[1], t.p[2]));
    }
    t.p[0] = b.lo, t.p[1] = b.hi, t.p[2] = b.lo;
    for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) sc.soup[9 * i + 3 * v + c] = t.p[v][c];
    if (sc.prim_quad.empty()) sc.prim_quad.assign(d.num_tris, 0);
    sc.quads.push_back(q);
    sc.prim_quad[i] = uint32_t(sc.quads.size());
}

int quadric_test(void *user, uint32_t prim, const float o[3], const float d[3], float t_max, float *t,
                 float b3[3]) {
    const Scene &sc = *static_cast<const Scene *>(user);
    quad::QuadricHit h;
    if (!quad::quadric_intersect(sc.quads[sc.prim_quad[prim] - 1], V3{o[0], o[1], o[2]}, V3{d[0], d[1], d[2]}, t_max,
                                 &h, kPi))
        return 0;
    *t = h.t;
    b3[0] = h.p.x, b3[1] = h.p.y, b3[2] = h.p.z;
    return 1;
}

void build_scene(Scene &sc, const orc_scene &d) {
    sc.desc = &d;
    sc.sp = &spectra();
    sc.cam = make_camera(d);
    const float *metal_data[8] = {pbrtm_data::kAgEta, pbrtm_data::kAgK, pbrtm_data::kAlEta, pbrtm_data::kAlK,
                                  pbrtm_data::kAuEta, pbrtm_data::kAuK, pbrtm_data::kCuEta, pbrtm_data::kCuK};
    const int metal_len[8] = {pbrtm_data::kAgEtaLen, pbrtm_data::kAgKLen, pbrtm_data::kAlEtaLen, pbrtm_data::kAlKLen,
                              pbrtm_data::kAuEtaLen, pbrtm_data::kAuKLen, pbrtm_data::kCuEtaLen, pbrtm_data::kCuKLen};
    for (int i = 0; i < 8; ++i) sc.metals[i] = pl_from_interleaved(metal_data[i], metal_len[i]);
    sc.glass_bk7 = pl_from_interleaved(pbrtm_data::kGlassBK7Eta, pbrtm_data::kGlassBK7EtaLen);
    // Materials (materials/diffuse_material.cu, materials/conductor_material.cu).
    for (uint32_t i = 0; i < d.num_materials; ++i) {
        const orc_material &m = d.materials[i];
        Material mm{};
        mm.type = m.type;
        if (m.type == 0) {
            if (m.rgb[0] > 1 || m.rgb[1] > 1 || m.rgb[2] > 1) throw std::runtime_error("albedo RGB > 1");
            float c[3] = {std::max(0.f, m.rgb[0]), std::max(0.f, m.rgb[1]), std::max(0.f, m.rgb[2])};
            mm.albedo = sc.sp->table.lookup(c);
        } else if (m.type == 1) {
            if (m.metal > 3) throw std::runtime_error("unknown metal");
            mm.eta = &sc.metals[2 * m.metal];
            mm.k = &sc.metals[2 * m.metal + 1];
            float u = m.uroughness, v = m.vroughness;
            if (m.remap_roughness) {
                u = std::sqrt(u);
                v = std::sqrt(v);
            }
            mm.ax = u;
            mm.ay = v;
        } else if (m.type == 2) {
            mm.eta_c = m.eta;
            mm.eta_pl = m.eta_named == 1 ? &sc.glass_bk7 : nullptr;
            if (m.eta_named > 1) throw std::runtime_error("unknown named eta");
            float u = m.uroughness, v = m.vroughness;
            if (m.remap_roughness) {
                u = std::sqrt(u);
                v = std::sqrt(v);
            }
            mm.ax = u;
            mm.ay = v;
        } else {
            throw std::runtime_error("unknown material type");
        }
        sc.mats.push_back(mm);
    }
    // Geometry in render space (triangle_mesh.cu:39-57: render_from_object applied per point).
    const M4 &rw = sc.cam.render_from_world.m;
    // render_from_object = render_from_world * CTM (scene_builder.cu get_render_from_object)
    std::vector<M4> rfo(d.num_xforms);
    for (uint32_t k = 0; k < d.num_xforms; ++k) {
        M4 x;
        std::memcpy(x.m, d.xforms + 16 * size_t(k), sizeof(x.m));
        rfo[k] = m4_mul(rw, x);
    }
    sc.tris.resize(d.num_tris);
    sc.soup.resize(d.num_tris * 9);
    for (uint64_t i = 0; i < d.num_tris; ++i) {
        Tri &t = sc.tris[i];
        const uint32_t xk = d.tri_xform ? d.tri_xform[i] : 0xFFFFFFFFu;
        if (xk != 0xFFFFFFFFu && xk >= d.num_xforms) throw std::runtime_error("transform index out of range");
        const M4 &m = xk == 0xFFFFFFFFu ? rw : rfo[xk];
        for (int v = 0; v < 3; ++v) {
            t.p[v] = xf_point(m, {d.tri[9 * i + 3 * v], d.tri[9 * i + 3 * v + 1], d.tri[9 * i + 3 * v + 2]});
            sc.soup[9 * i + 3 * v] = t.p[v].x;
            sc.soup[9 * i + 3 * v + 1] = t.p[v].y;
            sc.soup[9 * i + 3 * v + 2] = t.p[v].z;
        }
        t.material = d.tri_material[i];
        if (t.material >= d.num_materials) throw std::runtime_error("material index out of range");
        t.light = d.tri_light ? d.tri_light[i] : -1;
        t.flip = d.tri_reverse ? d.tri_reverse[i] != 0 : false;
        if (d.tri_shape && d.tri_shape[i] != 0xFFFFFFFFu) add_quadric(sc, d, i, sc.cam.render_from_world);
        if (t.light >= 0) {
            if (uint32_t(t.light) >= d.num_lights) throw std::runtime_error("light index out of range");
            t.light = int32_t(sc.lights.size());
            const orc_light &L = d.lights[d.tri_light[i]];
            Light l{};
            l.tri = uint32_t(i);
            l.two_sided = L.two_sided != 0;
            l.rgb_scale = float(2.0 * double(std::max(std::max(L.rgb[0], L.rgb[1]), L.rgb[2])));
            float c[3] = {0, 0, 0};
            if (l.rgb_scale > 0.0f)
                for (int q = 0; q < 3; ++q) c[q] = std::max(0.f, L.rgb[q] / l.rgb_scale);
            l.rsp = sc.sp->table.lookup(c);
            l.scale = L.scale / sc.sp->d65.inner(sc.sp->cie[1]);
            sc.lights.push_back(l);
        }
    }
    // Infinite lights join the light list at their place in file order (scene_builder.cu:494, :637).
    if (d.num_infinite_lights) {
        if (!d.infinite_lights) throw std::runtime_error("infinite lights missing");
        std::vector<Light> area = std::move(sc.lights), all;
        std::vector<int32_t> remap(area.size());
        for (size_t a = 0; a <= area.size(); ++a) {
            for (uint32_t q = 0; q < d.num_infinite_lights; ++q) {
                const orc_infinite_light &I = d.infinite_lights[q];
                const size_t at = std::min<size_t>(I.light_index, area.size());
                if (at != a) continue;
                // UniformInfiniteLight::create (uniform_infinite_light.cu:4-30)
                Light l{};
                l.infinite = true;
                l.two_sided = true;
                l.tri = 0xFFFFFFFFu;
                float scale = I.scale / sc.sp->d65.inner(sc.sp->cie[1]);
                if (I.illuminance > 0) scale *= I.illuminance / kPi;
                l.scale = scale;
                if (I.has_rgb) {
                    l.rgb_scale = float(2.0 * double(std::max(std::max(I.rgb[0], I.rgb[1]), I.rgb[2])));
                    float c[3] = {0, 0, 0};
                    if (l.rgb_scale > 0.0f)
                        for (int w = 0; w < 3; ++w) c[w] = std::max(0.f, I.rgb[w] / l.rgb_scale);
                    l.rsp = sc.sp->table.lookup(c);
                } else {
                    l.rgb_scale = 1;
                    l.rsp = Sigmoid{0, 0, std::numeric_limits<float>::infinity()};
                }
                sc.infinite.push_back(uint32_t(all.size()));
                all.push_back(l);
            }
            if (a < area.size()) {
                remap[a] = int32_t(all.size());
                all.push_back(area[a]);
            }
        }
        sc.lights = std::move(all);
        for (Tri &t : sc.tris)
            if (t.light >= 0) t.light = remap[t.light];
        // Bounds3::bounding_sphere of the aggregate's bounds (uniform_infinite_light.h:17-19)
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t i = 0; i < sc.soup.size(); i += 3)
            for (int c = 0; c < 3; ++c) lo[c] = std::min(lo[c], sc.soup[i + c]), hi[c] = std::max(hi[c], sc.soup[i + c]);
        const V3 ctr{(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
        const bool inside = ctr.x >= lo[0] && ctr.x <= hi[0] && ctr.y >= lo[1] && ctr.y <= hi[1] && ctr.z >= lo[2] &&
                            ctr.z <= hi[2];
        const float radius = inside ? length(sub(ctr, V3{hi[0], hi[1], hi[2]})) : 0.0f;
        sc.inf_dist = 2 * radius;
    }
    orc_build_info info;
    if (orc_bvh_build(sc.soup.data(), d.num_tris, &sc.bvh, &info) != 0) throw std::runtime_error("oracle BVH build");
    if (!sc.quads.empty()) {
        sc.is_quad.assign(d.num_tris, 0);
        for (uint64_t i = 0; i < d.num_tris; ++i) sc.is_quad[i] = sc.prim_quad[i] != 0;
        orc_bvh_set_custom(sc.bvh, &quadric_test, &sc, sc.is_quad.data());
    }
    // PowerLightSampler::create (power_light_sampler.cu:7-45) + Distribution1D::build.
    if (!sc.lights.empty()) {
        Lambda lam = sample_visible(0.5f);
        S4 pdf;
        for (int i = 0; i < kNS; ++i) pdf.v[i] = lam.pdf[i];
        std::vector<float> w;
        for (const Light &l : sc.lights) {
            S4 Ls;
            for (int i = 0; i < kNS; ++i) Ls.v[i] = l.rgb_scale * l.rsp(lam.l[i]);
            S4 phi;
            if (l.infinite) {  // UniformInfiniteLight::phi: 4 pi^2 R^2 scale L(lambda)
                const float R = sc.inf_dist / 2;
                phi = (sample_pl(sc.sp->d65, lam) * Ls) * (4 * kPi * kPi * (R * R) * l.scale);
            } else {
                S4 L = (sample_pl(sc.sp->d65, lam) * Ls) * l.scale;
                const float area = sc.prim_quad.empty() || !sc.prim_quad[l.tri]
                                       ? tri_area(sc.tris[l.tri])
                                       : quad::quadric_area(sc.quads[sc.prim_quad[l.tri] - 1]);
                float k = kPi * float(l.two_sided ? 2 : 1) * area;
                phi = L * k;
            }
            w.push_back(average(safe_div(phi, pdf)));
        }
        double sum = 0.0;
        for (float x : w) sum += x;
        size_t n = w.size();
        sc.light_pmf.resize(n);
        for (size_t i = 0; i < n; ++i) sc.light_pmf[i] = sum == 0.0 ? float(1.0 / n) : float(w[i] / sum);
        sc.light_cdf.resize(n);
        sc.light_cdf[0] = sc.light_pmf[0];
        for (size_t i = 1; i < n; ++i) sc.light_cdf[i] = sc.light_cdf[i - 1] + sc.light_pmf[i];
        sc.select_cdf = sc.light_cdf;
        if (d.select_scale) {
            if (!sc.infinite.empty()) throw std::runtime_error("select_scale with infinite lights");
            double ssum = 0.0;
            std::vector<double> ws(n);
            for (size_t i = 0; i < n; ++i) {
                const uint32_t g = uint32_t(d.tri_light[sc.lights[i].tri]);
                ws[i] = double(w[i]) * double(d.select_scale[g]) / double(d.lights[g].scale);
                ssum += ws[i];
            }
            float acc = 0;
            for (size_t i = 0; i < n; ++i) sc.select_cdf[i] = acc = acc + float(ws[i] / ssum);
        }
    }
    sc.out_from_sensor = sc.sp->rgb_from_xyz;  // times the identity white balance
}

// distribution_1d.h:9-42
inline uint32_t search_cdf(float u, const float *cdf, uint32_t len) {
    if (u < cdf[0] || len == 1) return 0;
    uint32_t start = 1, end = len;
    while (true) {
        if (end - start <= 10) {
            uint32_t t = start;
            for (uint32_t i = start; i < end; ++i)
                if (u >= cdf[i - 1] && u < cdf[i]) {
                    t = i;
                    break;
                }
            return t;
        }
        uint32_t mid = (start + end) / 2;
        if (u >= cdf[mid]) start = mid + 1;
        else end = mid + 1;
    }
}

// ------------------------------------------------------------ surface interaction
struct Isect {
    P3fi pi;
    V3 n, ns, dpdu, wo;
    uint32_t tri;
};
inline V3 p3fi_mid(const P3fi &p) { return p3fi_point(p); }

// shapes/triangle.cu:330-470 for meshes without normals or uv.
inline Isect make_isect(const Scene &sc, uint32_t prim, const TriHit &h, V3 ray_d) {
    if (!sc.prim_quad.empty() && sc.prim_quad[prim]) {
        const quad::QSurf q = quad::quadric_surface(sc.quads[sc.prim_quad[prim] - 1], V3{h.b0, h.b1, h.b2});
        Isect is;
        is.pi = q.pi;
        is.wo = normalize(neg(ray_d));
        is.n = q.n;
        is.ns = q.ns;
        is.dpdu = q.dpdu;
        is.tri = prim;
        return is;
    }
    const Tri &t = sc.tris[prim];
    V3 p0 = t.p[0], p1 = t.p[1], p2 = t.p[2];
    // uv = (0,0), (1,0), (1,1): duv02 = (-1,-1), duv12 = (0,-1), determinant 1.
    V3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
    const float du02[2] = {-1.f, -1.f}, du12[2] = {0.f, -1.f};
    float det = dop(du02[0], du12[1], du02[1], du12[0]);
    V3 dpdu, dpdv;
    bool degenerate = std::fabs(det) < 1e-9f;
    if (!degenerate) {
        float invdet = 1 / det;
        dpdu = scale(invdet, V3{dop(du12[1], dp02.x, du02[1], dp12.x), dop(du12[1], dp02.y, du02[1], dp12.y),
                                dop(du12[1], dp02.z, du02[1], dp12.z)});
        dpdv = scale(invdet, V3{dop(du02[0], dp12.x, du12[0], dp02.x), dop(du02[0], dp12.y, du12[0], dp02.y),
                                dop(du02[0], dp12.z, du12[0], dp02.z)});
    }
    if (degenerate || sqlen(cross(dpdu, dpdv)) == 0) {
        V3 ng = cross(sub(p2, p0), sub(p1, p0));
        coordinate_system(normalize(ng), &dpdu, &dpdv);
    }
    HitGeom g = hit_geometry(h, p0, p1, p2);
    Isect is;
    is.pi = g.pi;
    is.wo = normalize(neg(ray_d));
    is.n = normalize(cross(dp02, dp12));
    if (t.flip) is.n = neg(is.n);
    is.ns = is.n;
    is.dpdu = dpdu;
    is.tri = prim;
    return is;
}

// ------------------------------------------------------------ BSDFs
struct Frame {
    V3 x, y, z;
    V3 to_local(V3 v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    V3 from_local(V3 v) const { return add(add(scale(v.x, x), scale(v.y, y)), scale(v.z, z)); }
};
const float kInvPi = float(1.0 / double(kPi));

struct Cx {
    float re, im;
};
inline Cx cx(float r) { return {r, 0.f}; }
inline Cx operator+(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx operator-(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
inline Cx operator*(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx operator/(Cx a, Cx b) {
    float s = 1 / (b.re * b.re + b.im * b.im);
    return {s * (a.re * b.re + a.im * b.im), s * (a.im * b.re - a.re * b.im)};
}
inline float cnorm(Cx a) { return a.re * a.re + a.im * a.im; }
inline Cx csqrt(Cx z) {  // util/complex.h
    float n = std::sqrt(cnorm(z));
    float t1 = std::sqrt(.5f * (n + std::fabs(z.re)));
    float t2 = .5f * z.im / t1;
    if (n == 0) return cx(0);
    if (z.re >= 0) return {t1, t2};
    return {std::fabs(t2), std::copysign(t1, z.im)};
}
// util/scattering.h:56-72
inline float fr_complex(float cos_i, Cx eta) {
    cos_i = clampf(cos_i, 0, 1);
    float sin2_i = 1 - sqrf(cos_i);
    Cx sin2_t = cx(sin2_i) / (eta * eta);
    Cx cos_t = csqrt(cx(1) - sin2_t);
    Cx r_parl = (eta * cx(cos_i) - cos_t) / (eta * cx(cos_i) + cos_t);
    Cx r_perp = (cx(cos_i) - eta * cos_t) / (cx(cos_i) + eta * cos_t);
    return (cnorm(r_parl) + cnorm(r_perp)) / 2;
}
// util/scattering.h:82-102
inline float fr_dielectric(float cos_i, float eta) {
    cos_i = clampf(cos_i, -1, 1);
    if (cos_i < 0) {
        eta = 1 / eta;
        cos_i = -cos_i;
    }
    float sin2_i = 1 - sqrf(cos_i);
    float sin2_t = sin2_i / sqrf(eta);
    if (sin2_t >= 1) return 1.0f;
    float cos_t = safe_sqrt(1 - sin2_t);
    float r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (sqrf(r_parl) + sqrf(r_perp)) / 2;
}
// util/scattering.h:15-44
inline bool refract(V3 wi, V3 n, float eta, float *etap, V3 *wt) {
    float cos_i = dot(n, wi);
    if (cos_i < 0) {
        eta = 1 / eta;
        cos_i = -cos_i;
        n = neg(n);
    }
    float sin2_i = std::max(0.0f, 1 - sqrf(cos_i));
    float sin2_t = sin2_i / sqrf(eta);
    if (sin2_t >= 1) return false;
    float cos_t = std::sqrt(1 - sin2_t);
    *wt = add(divs(neg(wi), eta), scale(cos_i / eta - cos_t, n));
    *etap = eta;
    return true;
}
inline V3 reflect(V3 wo, V3 n) { return add(neg(wo), scale(2 * dot(wo, n), n)); }

// Local-frame trigonometry (euclidean_space/vector3.h:173-209).
inline float cos2_theta(V3 w) { return sqrf(w.z); }
inline float sin2_theta(V3 w) { return std::max(0.0f, float(1.0 - double(cos2_theta(w)))); }
inline float tan2_theta(V3 w) { return sin2_theta(w) / cos2_theta(w); }
inline float sin_theta(V3 w) { return std::sqrt(sin2_theta(w)); }
inline float cos_phi(V3 w) {
    float s = sin_theta(w);
    return (s == 0) ? 1 : clampf(w.x / s, -1, 1);
}
inline float sin_phi(V3 w) {
    float s = sin_theta(w);
    return (s == 0) ? 0 : clampf(w.y / s, -1, 1);
}
inline bool same_hemisphere(V3 a, V3 b) { return a.z * b.z > 0; }

// Trowbridge-Reitz microfacet distribution (util/scattering.h:105-229).
struct TR {
    float ax, ay;
    bool smooth() const { return std::max(ax, ay) < 1e-3f; }
    float D(V3 wm) const {
        float t2 = tan2_theta(wm);
        if (std::isinf(t2)) return 0;
        float cos4 = sqrf(cos2_theta(wm));
        if (cos4 < 1e-16f) return 0;
        float e = t2 * (sqrf(cos_phi(wm) / ax) + sqrf(sin_phi(wm) / ay));
        return 1 / (kPi * ax * ay * cos4 * sqrf(1 + e));
    }
    float lambda(V3 w) const {
        float t2 = tan2_theta(w);
        if (std::isinf(t2)) return 0;
        float a2 = sqrf(cos_phi(w) * ax) + sqrf(sin_phi(w) * ay);
        return (std::sqrt(1 + a2 * t2) - 1) / 2;
    }
    float G1(V3 w) const { return 1 / (1 + lambda(w)); }
    float G(V3 wo, V3 wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    float D(V3 w, V3 wm) const { return G1(w) / std::fabs(w.z) * D(wm) * abs_dot(w, wm); }
    float pdf(V3 w, V3 wm) const { return D(w, wm); }
    V3 sample_wm(V3 w, float u0, float u1) const {
        V3 wh = normalize({ax * w.x, ay * w.y, w.z});
        if (wh.z < 0) wh = neg(wh);
        V3 T1 = (wh.z < 0.99999f) ? normalize(cross({0, 0, 1}, wh)) : V3{1, 0, 0};
        V3 T2 = cross(wh, T1);
        float r = std::sqrt(u0), th = 2 * kPi * u1;  // sampling.h:85-90, polar disk
        float px = r * tcos(th), py = r * tsin(th);
        float h = std::sqrt(1 - sqrf(px));
        py = lerpf((1 + wh.z) / 2, h, py);
        float pz = std::sqrt(std::max(0.0f, 1 - (px * px + py * py)));
        V3 nh = add(add(scale(px, T1), scale(py, T2)), scale(pz, wh));
        return normalize({ax * nh.x, ay * nh.y, std::max(1e-6f, nh.z)});
    }
};

enum : uint32_t { kDiffuse = 0, kConductor = 1, kDielectric = 2 };
// BxDFFlags (base/bxdf.h): reflection 1, transmission 2, diffuse 4, glossy 8, specular 16
struct Bsdf {
    Frame fr;
    uint32_t type;
    S4 r;          // diffuse reflectance
    S4 eta, k;     // conductor
    float eta_d;   // dielectric
    TR tr;
    bool diffuse_set;
};

inline Bsdf make_bsdf(const Scene &sc, const Isect &is, Lambda &lam) {
    const Material &m = sc.mats[sc.tris[is.tri].material];
    Bsdf b;
    V3 x = normalize(is.dpdu);
    b.fr = {x, cross(is.ns, x), is.ns};
    b.type = m.type;
    b.tr = TR{m.ax, m.ay};
    if (!b.tr.smooth()) b.tr = TR{std::max(m.ax, 1e-4f), std::max(m.ay, 1e-4f)};
    if (m.type == kDiffuse) {
        S4 r = sample_sigmoid(m.albedo, lam);
        for (int i = 0; i < kNS; ++i) r.v[i] = clampf(r.v[i], 0.0f, 1.0f);
        b.r = r;
        b.diffuse_set = positive(r);
    } else if (m.type == kConductor) {
        b.eta = sample_pl(*m.eta, lam);
        b.k = sample_pl(*m.k, lam);
    } else {
        // materials/dielectric_material.cu:37-70
        float e = m.eta_pl ? (*m.eta_pl)(lam.l[0]) : m.eta_c;
        if (m.eta_pl) {  // terminate_secondary (sampled_wavelengths.h:84-95)
            bool done = true;
            for (int i = 1; i < kNS; ++i) done = done && lam.pdf[i] == 0;
            if (!done) {
                for (int i = 1; i < kNS; ++i) lam.pdf[i] = 0;
                lam.pdf[0] /= kNS;
            }
        }
        if (e == 0) e = 1;
        b.eta_d = e;
    }
    return b;
}
inline uint32_t bsdf_flags(const Bsdf &b) {
    if (b.type == kDiffuse) return b.diffuse_set ? (1u | 4u) : 0u;
    if (b.type == kConductor) return b.tr.smooth() ? (1u | 16u) : (1u | 8u);
    uint32_t f = b.eta_d == 1 ? 2u : (1u | 2u);
    return f | (b.tr.smooth() ? 16u : 8u);
}
inline bool non_specular(uint32_t f) { return (f & (4u | 8u)) != 0; }

// BxDF::f in the local frame; `valid` is false where the reference returns an
// unset (NaN) spectrum, which callers treat as zero.
inline S4 bxdf_f(const Bsdf &b, V3 wo, V3 wi) {
    if (b.type == kDiffuse) {
        if (!same_hemisphere(wo, wi)) return s4(0.f);
        return b.r * kInvPi;
    }
    if (b.type == kConductor) {  // bxdfs/conductor_bxdf.cu:54-83
        if (!same_hemisphere(wo, wi) || b.tr.smooth()) return s4(0.f);
        float co = std::fabs(wo.z), ci = std::fabs(wi.z);
        if (ci == 0 || co == 0) return s4(0.f);
        V3 wm = add(wi, wo);
        if (sqlen(wm) == 0) return s4(0.f);
        wm = normalize(wm);
        S4 F;
        for (int i = 0; i < kNS; ++i) F.v[i] = fr_complex(abs_dot(wo, wm), Cx{b.eta.v[i], b.k.v[i]});
        return ((F * b.tr.D(wm)) * b.tr.G(wo, wi)) / (4 * ci * co);
    }
    // bxdfs/dielectric_bxdf.cu:113-164
    if (b.eta_d == 1 || b.tr.smooth()) return s4(0.f);
    float co = wo.z, ci = wi.z;
    bool refl = ci * co > 0;
    float etap = 1;
    if (!refl) etap = co > 0 ? b.eta_d : (1 / b.eta_d);
    V3 wm = add(scale(etap, wi), wo);
    if (ci == 0 || co == 0 || sqlen(wm) == 0) return s4(0.f);
    wm = normalize(face_forward(wm, {0, 0, 1}));
    if (dot(wm, wi) * ci < 0 || dot(wm, wo) * co < 0) return s4(0.f);
    float F = fr_dielectric(dot(wo, wm), b.eta_d);
    if (refl) return s4(b.tr.D(wm) * b.tr.G(wo, wi) * F / std::fabs(4 * ci * co));
    float denom = sqrf(dot(wi, wm) + dot(wo, wm) / etap) * ci * co;
    float ft = b.tr.D(wm) * (1 - F) * b.tr.G(wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / denom);
    ft /= sqrf(etap);
    return s4(ft);
}
inline float bxdf_pdf(const Bsdf &b, V3 wo, V3 wi) {
    if (b.type == kDiffuse) {
        if (!same_hemisphere(wo, wi)) return 0;
        return std::fabs(wi.z) / kPi;
    }
    if (b.type == kConductor) {  // conductor_bxdf.cu:85-106
        if (!same_hemisphere(wo, wi) || b.tr.smooth()) return 0;
        V3 wm = add(wo, wi);
        if (sqlen(wm) == 0) return 0;
        wm = face_forward(normalize(wm), {0, 0, 1});
        return b.tr.pdf(wo, wm) / (4 * abs_dot(wo, wm));
    }
    // dielectric_bxdf.cu:166-230
    if (b.eta_d == 1 || b.tr.smooth()) return 0;
    float co = wo.z, ci = wi.z;
    bool refl = ci * co > 0;
    float etap = 1;
    if (!refl) etap = co > 0 ? b.eta_d : (1 / b.eta_d);
    V3 wm = add(scale(etap, wi), wo);
    if (ci == 0 || co == 0 || sqlen(wm) == 0) return 0;
    wm = normalize(face_forward(wm, {0, 0, 1}));
    if (dot(wm, wi) * ci < 0 || dot(wm, wo) * co < 0) return 0;
    float R = fr_dielectric(dot(wo, wm), b.eta_d);
    float T = 1 - R;
    float pr = R, pt = T;
    if (pr == 0 && pt == 0) return 0;
    if (refl) return b.tr.pdf(wo, wm) / (4 * abs_dot(wo, wm)) * pr / (pr + pt);
    float denom = sqrf(dot(wi, wm) + dot(wo, wm) / etap);
    float dwm_dwi = abs_dot(wi, wm) / denom;
    return b.tr.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
}
// BSDF::f / BSDF::pdf (base/bsdf.cu:44-93): render space in, local evaluation.
inline S4 bsdf_f(const Bsdf &b, V3 wo_r, V3 wi_r) {
    V3 wi = b.fr.to_local(wi_r), wo = b.fr.to_local(wo_r);
    if (wo.z == 0) return s4(0.f);
    return bxdf_f(b, wo, wi);
}
inline float bsdf_pdf(const Bsdf &b, V3 wo_r, V3 wi_r) {
    V3 wo = b.fr.to_local(wo_r), wi = b.fr.to_local(wi_r);
    if (wo.z == 0) return 0;
    return bxdf_pdf(b, wo, wi);
}
struct BS {
    S4 f;
    V3 wi;
    float pdf, eta;
    bool specular, transmission;
};
// BxDF::sample_f for each lobe, then BSDF::sample_f's checks (base/bsdf.cu:57-80).
inline bool bsdf_sample(const Bsdf &b, V3 wo_r, float uc, float u0, float u1, BS *out) {
    V3 wo = b.fr.to_local(wo_r);
    if (wo.z == 0 || bsdf_flags(b) == 0) return false;
    BS bs;
    bs.eta = 1;
    bs.transmission = false;
    if (b.type == kDiffuse) {
        V3 wi = sample_cosine_hemisphere(u0, u1);
        if (wo.z < 0) wi.z *= -1;
        bs.pdf = std::fabs(wi.z) / kPi;
        bs.f = b.r * kInvPi;
        bs.wi = wi;
        bs.specular = false;
    } else if (b.type == kConductor) {
        if (b.tr.smooth()) {
            V3 wi = {-wo.x, -wo.y, wo.z};
            float c = std::fabs(wi.z);
            S4 F;
            for (int i = 0; i < kNS; ++i) F.v[i] = fr_complex(c, Cx{b.eta.v[i], b.k.v[i]});
            bs.f = F / c;
            bs.wi = wi;
            bs.pdf = 1;
            bs.specular = true;
        } else {  // conductor_bxdf.cu:24-52
            if (wo.z == 0) return false;
            V3 wm = b.tr.sample_wm(wo, u0, u1);
            V3 wi = reflect(wo, wm);
            if (!same_hemisphere(wo, wi)) return false;
            float pdf = b.tr.pdf(wo, wm) / (4 * abs_dot(wo, wm));
            float co = std::fabs(wo.z), ci = std::fabs(wi.z);
            if (ci == 0 || co == 0) return false;
            S4 F;
            for (int i = 0; i < kNS; ++i) F.v[i] = fr_complex(abs_dot(wo, wm), Cx{b.eta.v[i], b.k.v[i]});
            bs.f = ((F * b.tr.D(wm)) * b.tr.G(wo, wi)) / (4 * ci * co);
            bs.wi = wi;
            bs.pdf = pdf;
            bs.specular = false;
        }
    } else if (b.eta_d == 1 || b.tr.smooth()) {  // dielectric_bxdf.cu:4-49
        float R = fr_dielectric(wo.z, b.eta_d), T = 1 - R;
        float pr = R, pt = T;
        if (pr == 0 && pt == 0) return false;
        if (uc < pr / (pr + pt)) {
            V3 wi = {-wo.x, -wo.y, wo.z};
            bs.f = s4(R / std::fabs(wi.z));
            bs.wi = wi;
            bs.pdf = pr / (pr + pt);
            bs.specular = true;
        } else {
            V3 wi;
            float etap;
            if (!refract(wo, {0, 0, 1}, b.eta_d, &etap, &wi)) return false;
            S4 ft = s4(T / std::fabs(wi.z));
            ft = ft / sqrf(etap);
            bs.f = ft;
            bs.wi = wi;
            bs.pdf = pt / (pr + pt);
            bs.specular = true;
            bs.transmission = true;
            bs.eta = etap;
        }
    } else {  // rough dielectric (dielectric_bxdf.cu:50-110)
        V3 wm = b.tr.sample_wm(wo, u0, u1);
        float R = fr_dielectric(dot(wo, wm), b.eta_d), T = 1 - R;
        float pr = R, pt = T;
        if (pr == 0 && pt == 0) return false;
        if (uc < pr / (pr + pt)) {
            V3 wi = reflect(wo, wm);
            if (!same_hemisphere(wo, wi)) return false;
            bs.pdf = b.tr.pdf(wo, wm) / (4 * abs_dot(wo, wm)) * pr / (pr + pt);
            bs.f = s4(b.tr.D(wm) * b.tr.G(wo, wi) * R / (4 * wi.z * wo.z));
            bs.wi = wi;
            bs.specular = false;
        } else {
            float etap;
            V3 wi;
            bool tir = !refract(wo, wm, b.eta_d, &etap, &wi);
            if (same_hemisphere(wo, wi) || wi.z == 0 || tir) return false;
            float denom = sqrf(dot(wi, wm) + dot(wo, wm) / etap);
            float dwm_dwi = abs_dot(wi, wm) / denom;
            bs.pdf = b.tr.pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
            float ft = T * b.tr.D(wm) * b.tr.G(wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom));
            S4 f = s4(ft);
            f = f / sqrf(etap);
            bs.f = f;
            bs.wi = wi;
            bs.specular = false;
            bs.transmission = true;
            bs.eta = etap;
        }
    }
    if (!positive(bs.f) || bs.pdf == 0 || bs.wi.z == 0) return false;
    bs.wi = b.fr.from_local(bs.wi);
    *out = bs;
    return true;
}

// ------------------------------------------------------------ triangle light sampling
inline float spherical_triangle_area(V3 a, V3 b, V3 c) {
    return float(std::fabs(2.0 * double(tatan2(dot(a, cross(b, c)), 1 + dot(a, b) + dot(a, c) + dot(b, c)))));
}
inline float tri_solid_angle(const Tri &t, V3 p) {
    return spherical_triangle_area(normalize(sub(t.p[0], p)), normalize(sub(t.p[1], p)), normalize(sub(t.p[2], p)));
}
inline float sample_linear(float u, float a, float b) {
    if (u == 0 && a == 0) return 0;
    float x = u * (a + b) / (a + std::sqrt(lerpf(u, sqrf(a), sqrf(b))));
    return std::min(x, kOneMinusEps);
}
inline void sample_bilinear(float u0, float u1, const float w[4], float *p0, float *p1) {
    float y = sample_linear(u1, w[0] + w[1], w[2] + w[3]);
    float x = sample_linear(u0, lerpf(y, w[0], w[2]), lerpf(y, w[1], w[3]));
    *p0 = x;
    *p1 = y;
}
inline float bilinear_pdf(float p0, float p1, const float w[4]) {
    if (p0 < 0 || p0 > 1 || p1 < 0 || p1 > 1) return 0;
    if (w[0] + w[1] + w[2] + w[3] == 0) return 1;
    float s = (1 - p0) * (1 - p1) * w[0] + p0 * (1 - p1) * w[1] + (1 - p0) * p1 * w[2] + p0 * p1 * w[3];
    return float(4.0 * double(s) / double(w[0] + w[1] + w[2] + w[3]));
}
// util/sampling.h:109-201
inline void sample_spherical_triangle(float out[3], float *pdf, const V3 v[3], V3 p, float u0, float u1) {
    *pdf = 0;
    V3 a = normalize(sub(v[0], p)), b = normalize(sub(v[1], p)), c = normalize(sub(v[2], p));
    V3 n_ab = cross(a, b), n_bc = cross(b, c), n_ca = cross(c, a);
    if (sqlen(n_ab) == 0 || sqlen(n_bc) == 0 || sqlen(n_ca) == 0) return;
    n_ab = normalize(n_ab);
    n_bc = normalize(n_bc);
    n_ca = normalize(n_ca);
    float alpha = angle_between(n_ab, neg(n_ca));
    float beta = angle_between(n_bc, neg(n_ab));
    float gamma = angle_between(n_ca, neg(n_bc));
    float A_pi = alpha + beta + gamma;
    float Ap_pi = lerpf(u0, kPi, A_pi);
    float A = A_pi - kPi;
    *pdf = (A <= 0) ? 0 : 1 / A;
    float cosAlpha = tcos(alpha), sinAlpha = tsin(alpha);
    float sinPhi = tsin(Ap_pi) * cosAlpha - tcos(Ap_pi) * sinAlpha;
    float cosPhi = tcos(Ap_pi) * cosAlpha + tsin(Ap_pi) * sinAlpha;
    float k1 = cosPhi + cosAlpha;
    float k2 = sinPhi - sinAlpha * dot(a, b);
    float cosBp = (k2 + (dop(k2, cosPhi, k1, sinPhi)) * cosAlpha) / ((sop(k2, sinPhi, k1, cosPhi)) * sinAlpha);
    cosBp = clampf(cosBp, -1, 1);
    float sinBp = safe_sqrt(1 - sqrf(cosBp));
    V3 cp = add(scale(cosBp, a), scale(sinBp, normalize(gram_schmidt(c, a))));
    float cosTheta = 1 - u1 * (1 - dot(cp, b));
    float sinTheta = safe_sqrt(1 - sqrf(cosTheta));
    V3 w = add(scale(cosTheta, b), scale(sinTheta, normalize(gram_schmidt(cp, b))));
    V3 e1 = sub(v[1], v[0]), e2 = sub(v[2], v[0]);
    V3 s1 = cross(w, e2);
    float divisor = dot(s1, e1);
    if (divisor == 0) {
        out[0] = out[1] = out[2] = float(1.0 / 3.0);
        return;
    }
    float inv = 1 / divisor;
    V3 s = sub(p, v[0]);
    float b1 = dot(s, s1) * inv;
    float b2 = dot(w, cross(s, e1)) * inv;
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
// util/sampling.cu:41-103 (Arvo's inversion)
inline void invert_spherical_triangle(const V3 v[3], V3 p, V3 w, float *u0, float *u1) {
    V3 a = normalize(sub(v[0], p)), b = normalize(sub(v[1], p)), c = normalize(sub(v[2], p));
    V3 n_ab = cross(a, b), n_bc = cross(b, c), n_ca = cross(c, a);
    if (sqlen(n_ab) == 0 || sqlen(n_bc) == 0 || sqlen(n_ca) == 0) {
        *u0 = *u1 = 0;
        return;
    }
    n_ab = normalize(n_ab);
    n_bc = normalize(n_bc);
    n_ca = normalize(n_ca);
    float alpha = angle_between(n_ab, neg(n_ca));
    float beta = angle_between(n_bc, neg(n_ab));
    float gamma = angle_between(n_ca, neg(n_bc));
    V3 cp = normalize(cross(cross(b, w), cross(c, a)));
    if (dot(cp, add(a, c)) < 0) cp = neg(cp);
    float uu0;
    if (dot(a, cp) > 0.99999847691f) {
        uu0 = 0;
    } else {
        V3 n_cpb = cross(cp, b), n_acp = cross(a, cp);
        if (sqlen(n_cpb) == 0 || sqlen(n_acp) == 0) {
            *u0 = *u1 = 0.5f;
            return;
        }
        n_cpb = normalize(n_cpb);
        n_acp = normalize(n_acp);
        float Ap = alpha + angle_between(n_ab, n_cpb) + angle_between(n_acp, neg(n_cpb)) - kPi;
        float A = alpha + beta + gamma - kPi;
        uu0 = Ap / A;
    }
    float uu1 = (1 - dot(w, b)) / (1 - dot(cp, b));
    *u0 = clampf(uu0, 0, 1);
    *u1 = clampf(uu1, 0, 1);
}

struct LightCtx {  // LightSampleContext / ShapeSampleContext
    P3fi pi;
    V3 n, ns;
};
struct ShapeSample {
    P3fi pi;
    V3 n;
    float pdf;
};
constexpr float kMinSA = 3e-4f, kMaxSA = 6.22f;

inline V3 light_normal(const Tri &t) {
    V3 n = normalize(cross(sub(t.p[1], t.p[0]), sub(t.p[2], t.p[0])));
    return t.flip ? neg(n) : n;
}
inline P3fi sample_point(const Tri &t, const float b[3], V3 *p_out) {
    V3 pabs = add(add(vabs(scale(b[0], t.p[0])), vabs(scale(b[1], t.p[1]))), vabs(scale(1 - b[0] - b[1], t.p[2])));
    V3 perr = scale(gammaf(6), pabs);
    V3 p = add(add(scale(b[0], t.p[0]), scale(b[1], t.p[1])), scale(b[2], t.p[2]));
    *p_out = p;
    return make_p3fi(p, perr);
}
// shapes/triangle.cu:126-215 (with the area-sampling path of :78-124)
inline bool tri_sample(const Tri &t, const LightCtx &ctx, float u0, float u1, ShapeSample *ss) {
    V3 cp = p3fi_mid(ctx.pi);
    float sa = tri_solid_angle(t, cp);
    if (sa < kMinSA || sa > kMaxSA) {
        float b[3];
        if (u0 < u1) {
            b[0] = u0 / 2;
            b[1] = u1 - b[0];
        } else {
            b[1] = u1 / 2;
            b[0] = u0 - b[1];
        }
        b[2] = 1 - b[0] - b[1];
        V3 p;
        ss->pi = sample_point(t, b, &p);
        ss->n = light_normal(t);
        ss->pdf = float(1.0) / tri_area(t);
        V3 pp = p3fi_mid(ss->pi);
        V3 wi = sub(pp, cp);
        if (sqlen(wi) == 0) return false;
        wi = normalize(wi);
        ss->pdf /= abs_dot(ss->n, neg(wi)) / sqlen(sub(cp, pp));
        if (std::isinf(ss->pdf)) return false;
        return true;
    }
    float pdf = 1.0f;
    if (!v3_zero(ctx.ns)) {
        V3 wi[3] = {normalize(sub(t.p[0], cp)), normalize(sub(t.p[1], cp)), normalize(sub(t.p[2], cp))};
        float w[4] = {std::max(0.01f, abs_dot(ctx.ns, wi[1])), std::max(0.01f, abs_dot(ctx.ns, wi[1])),
                      std::max(0.01f, abs_dot(ctx.ns, wi[0])), std::max(0.01f, abs_dot(ctx.ns, wi[2]))};
        float a0, a1;
        sample_bilinear(u0, u1, w, &a0, &a1);
        u0 = a0;
        u1 = a1;
        pdf = bilinear_pdf(u0, u1, w);
    }
    float b[3], tri_pdf;
    sample_spherical_triangle(b, &tri_pdf, t.p, cp, u0, u1);
    if (tri_pdf == 0.0f) return false;
    pdf *= tri_pdf;
    V3 p;
    ss->pi = sample_point(t, b, &p);
    ss->n = light_normal(t);
    ss->pdf = pdf;
    return true;
}
// ShapeSampleContext::offset_ray_origin (base/shape.h:31-53): |n| . err, not |n . err|.
inline V3 ctx_offset_origin(const LightCtx &c, V3 w) {
    V3 err = p3fi_error(c.pi);
    float d = dot(vabs(c.n), err);
    V3 off = scale(d, c.n);
    if (dot(c.n, w) < 0) off = neg(off);
    V3 po = add(p3fi_mid(c.pi), off);
    for (int i = 0; i < 3; ++i) {
        if (off[i] > 0) po[i] = next_float_up(po[i]);
        else if (off[i] < 0) po[i] = next_float_down(po[i]);
    }
    return po;
}
// shapes/triangle.cu:20-76
inline float tri_pdf(const Tri &t, const LightCtx &ctx, V3 wi) {
    V3 cp = p3fi_mid(ctx.pi);
    float sa = tri_solid_angle(t, cp);
    if (sa < kMinSA || sa > kMaxSA) {
        V3 o = ctx_offset_origin(ctx, wi);
        TriHit h;
        if (!intersect_triangle(o, wi, std::numeric_limits<float>::infinity(), t.p[0], t.p[1], t.p[2], &h)) return 0;
        HitGeom g = hit_geometry(h, t.p[0], t.p[1], t.p[2]);
        V3 n = g.n;
        if (t.flip) n = neg(n);
        V3 ip = p3fi_mid(g.pi);
        float pdf = (1 / tri_area(t)) / (abs_dot(n, neg(wi)) / sqlen(sub(cp, ip)));
        if (std::isinf(pdf)) pdf = 0;
        return pdf;
    }
    float pdf = 1 / sa;
    if (!v3_zero(ctx.ns)) {
        float u0, u1;
        invert_spherical_triangle(t.p, cp, wi, &u0, &u1);
        V3 w3[3] = {normalize(sub(t.p[0], cp)), normalize(sub(t.p[1], cp)), normalize(sub(t.p[2], cp))};
        float w[4] = {std::max(0.01f, abs_dot(ctx.ns, w3[1])), std::max(0.01f, abs_dot(ctx.ns, w3[1])),
                      std::max(0.01f, abs_dot(ctx.ns, w3[0])), std::max(0.01f, abs_dot(ctx.ns, w3[2]))};
        pdf *= bilinear_pdf(u0, u1, w);
    }
    return pdf;
}

// ------------------------------------------------------------ sphere and disk emitters
// Sphere::sample(u) (shapes/sphere.cu:152-181): uniform area sampling.
inline void sphere_sample_area(const quad::Quadric &s, float u0, float u1, ShapeSample *ss) {
    const float z = 1 - 2 * u0;  // sample_uniform_sphere (util/sampling.h:203-211)
    const float r = safe_sqrt(1 - sqrf(z));
    const float phi = 2 * kPi * u1;
    const V3 w = {r * tcos(phi), r * tsin(phi), z};
    V3 pObj = add(V3{0, 0, 0}, scale(s.radius, w));
    pObj = scale(s.radius / std::sqrt(sqrf(pObj.x - 0) + sqrf(pObj.y - 0) + sqrf(pObj.z - 0)), pObj);
    const V3 pObjError = scale(gammaf(5), vabs(pObj));
    V3 n = normalize(quad::tf_normal(s.rfo.inv, pObj));
    if (s.reverse) n = scale(-1.f, n);
    ss->pi = quad::to_p3fi(quad::tf_point(s.rfo.m, quad::iv3_of(pObj, pObjError)));
    ss->n = n;
    ss->pdf = float(1) / quad::sphere_area(s);
}
inline float sqdist(V3 a, V3 b) { return sqrf(a.x - b.x) + sqrf(a.y - b.y) + sqrf(a.z - b.z); }
// Sphere::sample(ctx, u) (shapes/sphere.cu:183-255).
inline bool sphere_sample(const quad::Quadric &s, const LightCtx &ctx, float u0, float u1, ShapeSample *ss) {
    const V3 pCenter = quad::tf_point(s.rfo.m, V3{0, 0, 0});
    const V3 cp = p3fi_mid(ctx.pi);
    const V3 pOrigin = ctx_offset_origin(ctx, sub(pCenter, cp));
    if (sqdist(pOrigin, pCenter) <= sqrf(s.radius)) {
        sphere_sample_area(s, u0, u1, ss);
        const V3 sp = p3fi_mid(ss->pi);
        V3 wi = sub(sp, cp);
        if (sqlen(wi) == 0) return false;
        wi = normalize(wi);
        ss->pdf /= abs_dot(ss->n, neg(wi)) / sqdist(cp, sp);
        return !std::isinf(ss->pdf);
    }
    const float sinThetaMax = s.radius / std::sqrt(sqdist(cp, pCenter));
    const float sin2ThetaMax = sqrf(sinThetaMax);
    const float cosThetaMax = safe_sqrt(1 - sin2ThetaMax);
    float oneMinusCosThetaMax = 1 - cosThetaMax;
    float cosTheta = (cosThetaMax - 1) * u0 + 1;
    float sin2Theta = 1 - sqrf(cosTheta);
    if (sin2ThetaMax < 0.00068523f /* sin^2(1.5 deg) */) {
        sin2Theta = sin2ThetaMax * u0;
        cosTheta = std::sqrt(1 - sin2Theta);
        oneMinusCosThetaMax = sin2ThetaMax / 2;
    }
    const float cosAlpha = sin2Theta / sinThetaMax + cosTheta * safe_sqrt(1 - sin2Theta / sqrf(sinThetaMax));
    const float sinAlpha = safe_sqrt(1 - sqrf(cosAlpha));
    const float phi = u1 * 2 * kPi;
    const V3 w = {clampf(sinAlpha, -1, 1) * tcos(phi), clampf(sinAlpha, -1, 1) * tsin(phi), clampf(cosAlpha, -1, 1)};
    V3 fx, fy;
    const V3 fz = normalize(sub(pCenter, cp));
    coordinate_system(fz, &fx, &fy);  // Frame::from_z
    const V3 mw = neg(w);
    V3 n = add(add(scale(mw.x, fx), scale(mw.y, fy)), scale(mw.z, fz));
    const V3 p = add(pCenter, scale(s.radius, n));
    if (s.reverse) n = scale(-1.f, n);
    ss->pi = make_p3fi(p, scale(gammaf(5), vabs(p)));
    ss->n = n;
    ss->pdf = 1 / (2 * kPi * oneMinusCosThetaMax);
    return true;
}
// Disk::sample(u) and sample(ctx, u) (shapes/disk.cu:39-94).
inline bool disk_sample(const quad::Quadric &s, const LightCtx &ctx, float u0, float u1, ShapeSample *ss) {
    float dx = 0, dy = 0;  // sample_uniform_disk_concentric (util/sampling.h:61-83)
    const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    if (!(ox == 0 && oy == 0)) {
        const float pi_over_4 = float(kPi / 4.0);
        float r, theta;
        if (std::fabs(ox) > std::fabs(oy)) {
            r = ox;
            theta = pi_over_4 * (oy / ox);
        } else {
            r = oy;
            theta = float(kPi / 2.0) - pi_over_4 * (ox / oy);
        }
        dx = r * tcos(theta);
        dy = r * tsin(theta);
    }
    const V3 pObj = {dx * s.radius, dy * s.radius, s.height};
    ss->pi = quad::to_p3fi(quad::tf_point(s.rfo.m, quad::iv3_exact(pObj)));
    V3 n = normalize(quad::tf_normal(s.rfo.inv, V3{0, 0, 1}));
    if (s.reverse) n = scale(-1.f, n);
    ss->n = n;
    ss->pdf = float(1.0 / quad::disk_area(s));
    const V3 cp = p3fi_mid(ctx.pi), sp = p3fi_mid(ss->pi);
    V3 wi = sub(sp, cp);
    if (sqlen(wi) == 0) return false;
    wi = normalize(wi);
    ss->pdf /= abs_dot(ss->n, neg(wi)) / sqdist(cp, sp);
    return !std::isinf(ss->pdf);
}
// Sphere::pdf (shapes/sphere.cu:257-297), Disk::pdf (shapes/disk.cu:20-37).
inline float quadric_pdf(const quad::Quadric &s, const LightCtx &ctx, V3 wi) {
    const V3 cp = p3fi_mid(ctx.pi);
    if (s.type == 1) {
        const V3 pCenter = quad::tf_point(s.rfo.m, V3{0, 0, 0});
        if (sqdist(ctx_offset_origin(ctx, sub(pCenter, cp)), pCenter) > sqrf(s.radius)) {
            const float sin2ThetaMax = s.radius * s.radius / sqdist(cp, pCenter);
            const float cosThetaMax = safe_sqrt(1 - sin2ThetaMax);
            float oneMinusCosThetaMax = 1 - cosThetaMax;
            if (sin2ThetaMax < 0.00068523f) oneMinusCosThetaMax = sin2ThetaMax / 2;
            return float(1.0 / (2.0 * kPi * oneMinusCosThetaMax));
        }
    }
    quad::QuadricHit h;
    const V3 o = ctx_offset_origin(ctx, wi);
    if (!quad::quadric_intersect(s, o, wi, std::numeric_limits<float>::infinity(), &h, kPi)) return 0;
    const quad::QSurf g = quad::quadric_surface(s, h.p);
    float pdf = float((1.0 / quad::quadric_area(s)) / (abs_dot(g.n, neg(wi)) / sqdist(cp, p3fi_mid(g.pi))));
    if (std::isinf(pdf)) pdf = 0;
    return pdf;
}
// DiffuseAreaLight::sample_li / pdf_li through the emitter's shape (diffuse_area_light.cu:48-74).
inline bool emitter_sample(const Scene &sc, uint32_t prim, const LightCtx &ctx, float u0, float u1, ShapeSample *ss) {
    if (!sc.prim_quad.empty() && sc.prim_quad[prim]) {
        const quad::Quadric &q = sc.quads[sc.prim_quad[prim] - 1];
        return q.type == 1 ? sphere_sample(q, ctx, u0, u1, ss) : disk_sample(q, ctx, u0, u1, ss);
    }
    return tri_sample(sc.tris[prim], ctx, u0, u1, ss);
}
inline float emitter_pdf(const Scene &sc, uint32_t prim, const LightCtx &ctx, V3 wi) {
    if (!sc.prim_quad.empty() && sc.prim_quad[prim]) return quadric_pdf(sc.quads[sc.prim_quad[prim] - 1], ctx, wi);
    return tri_pdf(sc.tris[prim], ctx, wi);
}

inline float power_heuristic(float f, float g) {  // sampling.h:259-269 with nf = ng = 1
    float ff = 1 * f, gg = 1 * g;
    if (std::isinf(sqrf(ff))) return 1;
    return sqrf(ff) / (sqrf(ff) + sqrf(gg));
}

// UniformInfiniteLight::sample_li (lights/uniform_infinite_light.cu:51-60) with
// sample_uniform_sphere / uniform_sphere_pdf (util/sampling.h:203-216).
inline V3 sample_uniform_sphere(float u0, float u1) {
    const float z = 1 - 2 * u0;
    const float r = safe_sqrt(1 - sqrf(z));
    const float phi = 2 * kPi * u1;
    return {r * tcos(phi), r * tsin(phi), z};
}
const float kUniformSpherePdf = float(1.0 / (4.0 * double(kPi)));

// The infinite lights' emission on a path that ends with a positive beta (the
// wavefront's control_logic, wavefront_path.cu:140-160) or escapes (megakernel,
// megakernel_path.cu:58-75): pdf_li with an incomplete pdf allowed is 0, so the
// MIS weight of a non-specular bounce is power_heuristic(pdf_bsdf, 0).
inline S4 add_infinite(const Scene &sc, S4 L, const S4 &beta, const Lambda &lam, uint32_t depth, bool specular,
                       float pdf_bsdf) {
    for (uint32_t li : sc.infinite) {
        const S4 Le = light_L(sc, sc.lights[li], V3{0, 0, 0}, V3{0, 0, 0}, lam);
        if (depth == 0 || specular) L = L + beta * Le;
        else L = L + (beta * power_heuristic(pdf_bsdf, sc.light_pmf[li] * 0.0f)) * Le;
    }
    return L;
}

// Ray::spawn_ray_to (base/ray.h:45-51)
inline void spawn_ray_to(const P3fi &from, V3 n_from, const P3fi &to, V3 n_to, V3 *o, V3 *d) {
    V3 pf = offset_ray_origin(from, n_from, sub(p3fi_mid(to), p3fi_mid(from)));
    V3 pt = offset_ray_origin(to, n_to, sub(pf, p3fi_mid(to)));
    *o = pf;
    *d = sub(pt, pf);
}

// ------------------------------------------------------------ the wavefront renderer
struct PathSlot {
    bool active = false, free = true;
    uint32_t pixel = 0, sample = 0, depth = 0;
    Sampler smp;
    Lambda lam;
    S4 L, beta;
    float filter_w = 0;
    bool specular = true, any_non_specular = false;
    float pdf_bsdf = NAN, eta_scale = 1;
    LightCtx prev;
    V3 ro, rd;
    bool hit = false;
    uint32_t prim = 0;
    TriHit h;
    // deferred NEE: contribution and its two shadow rays
    bool nee = false;
    S4 nee_L;
    V3 so[2], sd[2];
    // gradient replay: dLoss/dL per wavelength, the path's final L, the pending
    // NEE's adjoint-weighted value / MIS weight / light group, and this phase's
    // records (reduced in slot order after each parallel phase)
    double g[4] = {0, 0, 0, 0};
    S4 Ltot;
    double nee_G = 0, nee_w = 0;
    int nee_group = -1;
    int rec_group = -1;
    double rec_a = 0, rec_k = 0;
    int rec_mat = -1;
    double rec_c[3] = {0, 0, 0};
};

// Gradient mode of render(). Forward pass: Lfin / pdffin receive each path's
// final L and wavelength pdfs. Replay: A holds the per-pixel adjoint of the
// sensor RGB (dLoss/d(rgb_sum / weight_sum) through rgb_from_xyz), and the
// replay accumulates d_coef (per material, sigmoid coefficients), a (per light
// group, scale-linear part) and k (light-pmf part).
struct GradIO {
    std::vector<S4> *Lfin = nullptr, *pdffin = nullptr;  // forward
    const std::vector<S4> *Ltot = nullptr, *pdf = nullptr;  // replay
    const std::vector<double> *A = nullptr;
    std::vector<double> d_coef, a;
    double k = 0;
    std::vector<double> d_rgb, d_scale;  // per material [3], per light group
    bool hold_selection = false;         // hold the light choice's 1/pmf (not differentiated)
    double M[9] = {};                    // out_from_sensor, for the adjoint
};

void parallel(size_t n, int threads, const std::function<void(size_t)> &fn) {
    if (threads <= 1 || n < 256) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&] {
            for (;;) {
                size_t b = next.fetch_add(1024);
                if (b >= n) break;
                size_t e = std::min(n, b + 1024);
                for (size_t i = b; i < e; ++i) fn(i);
            }
        });
    for (auto &t : ts) t.join();
}

struct FilmRec {
    uint32_t pixel, sample;
    float rgb[3], w;
};

// PixelSensor::to_sensor_rgb (films/pixel_sensor.h:33-40) with imaging ratio 1.
inline void sensor_rgb(const Scene &sc, const S4 &L, const Lambda &lam, float rgb[3]) {
    S4 pdf;
    for (int i = 0; i < kNS; ++i) pdf.v[i] = lam.pdf[i];
    S4 s = safe_div(L, pdf);
    for (int c = 0; c < 3; ++c) rgb[c] = 1.0f * average(sample_pl(sc.sp->cie[c], lam) * s);
}

// path_start / path_finish (nullable, [W*H*spp] by path id): the round a path's
// first ray is traced in and the round it is recorded in the film.
void render(const orc_scene &d, int threads, float *film_rgb, float *aux, uint64_t *stats,
            uint32_t *path_start = nullptr, uint32_t *path_finish = nullptr, GradIO *gio = nullptr) {
    Scene sc;
    build_scene(sc, d);
    const uint32_t W = d.width, H = d.height;
    const uint32_t spd = uint32_t(std::sqrt(double(d.spp)));
    if (spd * spd != d.spp || d.spp == 0) throw std::runtime_error("spp must be a square");
    const uint64_t total = uint64_t(W) * H * d.spp;
    const uint32_t P = d.pool_size;
    std::vector<PathSlot> pool(P);
    std::vector<float> rgb_sum(size_t(W) * H * 3, 0.f), w_sum(size_t(W) * H, 0.f);
    uint64_t next_id = 0, n_rays = 0, n_shadow = 0, n_done = 0, round = 0;
    uint32_t gen_round = 0;  // the round whose ray cast first traces the paths gen() starts
    const bool replay = gio && gio->A;
    if (replay && !sc.infinite.empty()) throw std::runtime_error("gradients with infinite lights");
    if (replay) {
        gio->d_coef.assign(size_t(d.num_materials) * 3, 0.0);
        gio->a.assign(d.num_lights, 0.0);
        gio->k = 0;
    }
    if (gio && gio->Lfin) {
        gio->Lfin->assign(total, s4(0));
        gio->pdffin->assign(total, s4(0));
    }
    auto group_of = [&](uint32_t light) { return int(d.tri_light[sc.lights[light].tri]); };
    auto dotg = [](const double *g, const S4 &c) {
        double v = 0;
        for (int i = 0; i < kNS; ++i) v += g[i] * double(c.v[i]);
        return v;
    };
    const float inf = std::numeric_limits<float>::infinity();

    // generate_new_path (wavefront_path.cu:265-308)
    auto gen = [&](std::vector<uint32_t> &slots) {
        std::vector<std::pair<uint32_t, uint64_t>> work;
        for (uint32_t s : slots) {
            uint64_t id = next_id++;
            if (id >= total) {
                pool[s].active = false;
                pool[s].free = false;
                continue;
            }
            work.push_back({s, id});
            if (path_start) path_start[id] = gen_round;
        }
        parallel(work.size(), threads, [&](size_t k) {
            PathSlot &p = pool[work[k].first];
            uint64_t id = work[k].second;
            p = PathSlot();
            p.active = true;
            p.free = false;
            p.pixel = uint32_t(id % (uint64_t(W) * H));
            p.sample = uint32_t(id / (uint64_t(W) * H));
            p.smp.spd = spd;
            p.smp.start(p.pixel, p.sample, 0);
            float u0, u1;
            p.smp.get_2d(&u0, &u1);
            float fx, fy, fw;
            sc.filter.sample(u0, u1, &fx, &fy, &fw);
            float px = (float(p.pixel % W) + fx) + 0.5f, py = (float(p.pixel / W) + fy) + 0.5f;
            p.filter_w = fw;
            p.lam = sample_visible(p.smp.get_1d());
            V3 pc = xf_point(sc.cam.camera_from_raster.m, {px, py, 0});
            V3 dc = normalize(pc);
            p.ro = xf_point(sc.cam.render_from_camera.m, {0, 0, 0});
            p.rd = xf_vector(sc.cam.render_from_camera.m, dc);
            p.L = s4(0);
            p.beta = s4(1);
            p.depth = 0;
            if (replay) {
                // dLoss/dL_i = w / 4 * sum_c A_c x_c(lambda_i) / pdf_i (sensor_rgb, film add_sample)
                const S4 &pf = (*gio->pdf)[id];
                const double *A = &(*gio->A)[3 * size_t(p.pixel)];
                for (int i = 0; i < kNS; ++i) {
                    double v = 0;
                    if (pf.v[i] != 0)
                        for (int c = 0; c < 3; ++c) v += A[c] * double(sc.sp->cie[c](p.lam.l[i])) / double(pf.v[i]);
                    p.g[i] = double(fw) * v / kNS;
                }
                p.Ltot = (*gio->Ltot)[id];
            }
        });
    };

    const auto t_start = std::chrono::steady_clock::now();
    std::vector<uint32_t> all(P);
    for (uint32_t i = 0; i < P; ++i) all[i] = i;
    gen(all);
    std::vector<uint32_t> active_slots;
    for (;;) {
        active_slots.clear();
        for (uint32_t i = 0; i < P; ++i)
            if (pool[i].active) active_slots.push_back(i);
        if (active_slots.empty()) break;
        // ray_cast + the shadow rays of the previous round's NEE
        parallel(active_slots.size(), threads, [&](size_t k) {
            PathSlot &p = pool[active_slots[k]];
            float t, b[3];
            uint32_t prim;
            int rc = orc_trace_ray(sc.bvh, &p.ro.x, &p.rd.x, inf, 0, &prim, &t, b);
            if (rc < 0) throw std::runtime_error("oracle traversal stack overflow");
            p.hit = rc == 1;
            p.prim = prim;
            p.h = TriHit{b[0], b[1], b[2], t};
            if (p.nee) {
                bool occluded = false;
                for (int r = 0; r < 2 && !occluded; ++r) {
                    uint32_t sp;
                    int sr = orc_trace_ray(sc.bvh, &p.so[r].x, &p.sd[r].x, 0.6f, 1, &sp, nullptr, nullptr);
                    occluded = sr == 1;
                }
                if (!occluded) {
                    p.L = p.L + p.nee_L;
                    if (replay) {
                        // NEE, C = beta f cos Le w(pmf p) / (pmf p), w = g^2 / (g^2 + b^2),
                        // g = pmf p. Both the MIS weight's pmf and the light choice's 1/pmf
                        // attached: dC/dpmf = C (1 - 2w) / pmf, so a += G 2 (1 - w),
                        // k += G (1 - 2w). hold_selection holds the choice's 1/pmf
                        // (dC/dpmf = 2 C (1 - w) / pmf): a += G (3 - 2w), k += 2 G (1 - w).
                        p.rec_group = p.nee_group;
                        if (!gio->hold_selection) {
                            p.rec_a = p.nee_G * 2.0 * (1.0 - p.nee_w);
                            p.rec_k = p.nee_G * (1.0 - 2.0 * p.nee_w);
                        } else {
                            p.rec_a = p.nee_G * (3.0 - 2.0 * p.nee_w);
                            p.rec_k = p.nee_G * 2.0 * (1.0 - p.nee_w);
                        }
                    }
                }
                p.nee = false;
            }
        });
        for (uint32_t s : active_slots) {
            n_rays++;
            PathSlot &p = pool[s];
            if (p.rec_group >= 0) {
                gio->a[p.rec_group] += p.rec_a;
                gio->k += p.rec_k;
                p.rec_group = -1;
            }
        }
        // control_logic (wavefront_path.cu:92-234)
        std::vector<FilmRec> recs;
        std::vector<uint32_t> freed, shade;
        for (uint32_t s : active_slots) {
            PathSlot &p = pool[s];
            bool terminate = !p.hit || p.depth > d.max_depth || !positive(p.beta);
            if (!terminate && p.depth > 8) {
                float u = p.smp.get_1d();
                S4 rr = p.beta * p.eta_scale;
                if (max_comp(rr) < 1) {
                    float q = clampf(1 - max_comp(rr), 0, 0.95f);
                    if (u < q) {
                        p.beta = s4(0.0f);
                        terminate = true;
                    } else {
                        p.beta = p.beta / (1 - q);
                    }
                }
            }
            if (terminate) {
                if (positive(p.beta) && !sc.infinite.empty())
                    p.L = add_infinite(sc, p.L, p.beta, p.lam, p.depth, p.specular, p.pdf_bsdf);
                FilmRec r;
                r.pixel = p.pixel;
                r.sample = p.sample;
                r.w = p.filter_w;
                sensor_rgb(sc, p.L, p.lam, r.rgb);
                if (path_finish) path_finish[uint64_t(p.sample) * W * H + p.pixel] = uint32_t(round);
                if (gio && gio->Lfin) {
                    const uint64_t id = uint64_t(p.sample) * W * H + p.pixel;
                    (*gio->Lfin)[id] = p.L;
                    for (int i = 0; i < kNS; ++i) (*gio->pdffin)[id].v[i] = p.lam.pdf[i];
                }
                recs.push_back(r);
                p.active = false;
                p.free = true;
                freed.push_back(s);
                n_done++;
                continue;
            }
            const Tri &t = sc.tris[p.prim];
            if (t.light >= 0) {
                const Light &L = sc.lights[t.light];
                Isect is = make_isect(sc, p.prim, p.h, p.rd);
                S4 Le = light_L(sc, L, is.n, neg(p.rd), p.lam);
                if (positive(Le)) {
                    if (p.depth == 0 || p.specular) {
                        p.L = p.L + p.beta * Le;
                        if (replay) gio->a[group_of(t.light)] += dotg(p.g, p.beta * Le);
                    } else {
                        float pdf_light = sc.light_pmf[t.light] * emitter_pdf(sc, p.prim, p.prev, p.rd);
                        float wl = power_heuristic(p.pdf_bsdf, pdf_light);
                        p.L = p.L + (p.beta * wl) * Le;
                        if (replay) {  // emission with MIS: a += G (2w - 1), k += -2 G (1 - w)
                            const double G = dotg(p.g, (p.beta * wl) * Le);
                            gio->a[group_of(t.light)] += G * (2.0 * wl - 1.0);
                            gio->k += -2.0 * G * (1.0 - wl);
                        }
                    }
                }
            }
            p.depth += 1;
            shade.push_back(s);
        }
        // write_frame_buffer: sorted by (pixel, sample), film add_sample (rgb_film.cu:85-102)
        std::sort(recs.begin(), recs.end(), [](const FilmRec &a, const FilmRec &b) {
            return a.pixel != b.pixel ? a.pixel < b.pixel : a.sample < b.sample;
        });
        for (const FilmRec &r : recs) {
            if (r.w == 0) continue;
            for (int c = 0; c < 3; ++c) rgb_sum[3 * r.pixel + c] += r.w * r.rgb[c];
            w_sum[r.pixel] += r.w;
        }
        gen_round = uint32_t(round + 1);
        gen(freed);
        // gpu_evaluate_material + sample_bsdf (wavefront_path.cu:310-431)
        parallel(shade.size(), threads, [&](size_t k) {
            PathSlot &p = pool[shade[k]];
            Isect is = make_isect(sc, p.prim, p.h, p.rd);
            Bsdf b = make_bsdf(sc, is, p.lam);
            uint32_t flags = bsdf_flags(b);
            if (replay && b.type == kDiffuse) {
                // every later contribution is linear in this vertex's reflectance R:
                // dLoss/dR_i = g_i (Ltot_i - L_i) / R_i, then R = sigmoid(c0 l^2 + c1 l + c2)
                const Material &m = sc.mats[sc.tris[p.prim].material];
                p.rec_mat = int(sc.tris[p.prim].material);
                for (int q = 0; q < 3; ++q) p.rec_c[q] = 0;
                for (int i = 0; i < kNS; ++i) {
                    if (b.r.v[i] <= 0 || b.r.v[i] >= 1) continue;  // clamped
                    const double l = p.lam.l[i];
                    const double x = (double(m.albedo.c0) * l + double(m.albedo.c1)) * l + double(m.albedo.c2);
                    const double ds = 0.5 / std::pow(1.0 + x * x, 1.5);
                    const double v = p.g[i] * (double(p.Ltot.v[i]) - double(p.L.v[i])) / double(b.r.v[i]) * ds;
                    p.rec_c[0] += v * l * l;
                    p.rec_c[1] += v * l;
                    p.rec_c[2] += v;
                }
            }
            if (non_specular(flags)) {
                // sample_ld (wavefront_path.cu:574-624)
                LightCtx ctx{is.pi, is.n, is.ns};
                bool refl = (flags & 1u) != 0, trans = (flags & 2u) != 0;
                if (refl && !trans) {
                    V3 po = offset_ray_origin(is.pi, is.n, is.wo);
                    ctx.pi = make_p3fi(po, {0, 0, 0});
                } else if (trans && !refl) {
                    V3 po = offset_ray_origin(is.pi, is.n, neg(is.wo));
                    ctx.pi = make_p3fi(po, {0, 0, 0});
                }
                float u = p.smp.get_1d();
                float ul0, ul1;
                p.smp.get_2d(&ul0, &ul1);
                if (!sc.lights.empty() && sc.lights[search_cdf(u, sc.select_cdf.data(), uint32_t(sc.lights.size()))].infinite) {
                    const uint32_t li = search_cdf(u, sc.select_cdf.data(), uint32_t(sc.lights.size()));
                    const V3 wi = sample_uniform_sphere(ul0, ul1);
                    const S4 Le = light_L(sc, sc.lights[li], V3{0, 0, 0}, V3{0, 0, 0}, p.lam);
                    if (positive(Le)) {
                        const S4 f = bsdf_f(b, is.wo, wi) * abs_dot(wi, is.ns);
                        if (positive(f)) {
                            const float pdf_light = sc.light_pmf[li] * kUniformSpherePdf;
                            const float wl = power_heuristic(pdf_light, bsdf_pdf(b, is.wo, wi));
                            p.nee = true;
                            p.nee_L = p.beta * (((Le * wl) * f) / pdf_light);
                            // the light point: ctx.p() + wi * 2R, an exact point without a normal
                            const P3fi pl = make_p3fi(add(p3fi_mid(ctx.pi), scale(sc.inf_dist, wi)), {0, 0, 0});
                            spawn_ray_to(is.pi, is.n, pl, V3{0, 0, 0}, &p.so[0], &p.sd[0]);
                            spawn_ray_to(pl, V3{0, 0, 0}, is.pi, is.n, &p.so[1], &p.sd[1]);
                        }
                    }
                } else if (!sc.lights.empty()) {
                    uint32_t li = search_cdf(u, sc.select_cdf.data(), uint32_t(sc.lights.size()));
                    float pmf = sc.light_pmf[li];
                    const Light &L = sc.lights[li];
                    ShapeSample ss;
                    bool ok = emitter_sample(sc, L.tri, ctx, ul0, ul1, &ss);
                    V3 cpt = p3fi_mid(ctx.pi);
                    V3 sp = p3fi_mid(ss.pi);
                    if (ok && !(ss.pdf == 0 || sqlen(sub(sp, cpt)) == 0)) {
                        V3 wi = normalize(sub(sp, cpt));
                        S4 Le = light_L(sc, L, ss.n, neg(wi), p.lam);
                        if (positive(Le) && ss.pdf != 0) {
                            S4 f = bsdf_f(b, is.wo, wi) * abs_dot(wi, is.ns);
                            if (positive(f)) {
                                float pdf_light = pmf * ss.pdf;
                                float pdf_b = bsdf_pdf(b, is.wo, wi);
                                float wl = power_heuristic(pdf_light, pdf_b);
                                S4 Ld = ((Le * wl) * f) / pdf_light;
                                p.nee = true;
                                p.nee_L = p.beta * Ld;
                                if (replay) {
                                    p.nee_G = dotg(p.g, p.nee_L);
                                    p.nee_w = wl;
                                    p.nee_group = group_of(li);
                                }
                                spawn_ray_to(is.pi, is.n, ss.pi, ss.n, &p.so[0], &p.sd[0]);
                                spawn_ray_to(ss.pi, ss.n, is.pi, is.n, &p.so[1], &p.sd[1]);
                            }
                        }
                    }
                }
            }
            V3 wo = neg(p.rd);
            float uc = p.smp.get_1d();
            float u0, u1;
            p.smp.get_2d(&u0, &u1);
            BS bs;
            if (!bsdf_sample(b, wo, uc, u0, u1, &bs)) {
                p.beta = s4(0.0f);
                return;
            }
            p.beta = p.beta * ((bs.f * abs_dot(bs.wi, is.ns)) / bs.pdf);
            p.pdf_bsdf = bs.pdf;
            p.specular = bs.specular;
            p.any_non_specular |= !bs.specular;
            if (bs.transmission) p.eta_scale *= sqrf(bs.eta);
            p.prev = LightCtx{is.pi, is.n, is.ns};
            p.ro = offset_ray_origin(is.pi, is.n, bs.wi);
            p.rd = bs.wi;
        });
        for (uint32_t s : shade) {
            if (pool[s].nee) n_shadow += 2;
            PathSlot &p = pool[s];
            if (p.rec_mat >= 0) {
                for (int q = 0; q < 3; ++q) gio->d_coef[3 * size_t(p.rec_mat) + q] += p.rec_c[q];
                p.rec_mat = -1;
            }
        }
        round++;
    }
    // RGBFilm::get_pixel_rgb (rgb_film.cu:129-144)
    for (uint32_t px = 0; px < W * H; ++px) {
        float rgb[3] = {rgb_sum[3 * px], rgb_sum[3 * px + 1], rgb_sum[3 * px + 2]};
        if (w_sum[px] != 0)
            for (int c = 0; c < 3; ++c) rgb[c] /= w_sum[px];
        auto o = m3_apply(sc.out_from_sensor, rgb);
        for (int c = 0; c < 3; ++c) film_rgb[3 * size_t(px) + c] = o[c];
        if (aux) {
            for (int c = 0; c < 3; ++c) aux[4 * size_t(px) + c] = rgb_sum[3 * px + c];
            aux[4 * size_t(px) + 3] = w_sum[px];
        }
    }
    if (gio) {
        for (int q = 0; q < 9; ++q) gio->M[q] = double(sc.out_from_sensor.m[q / 3][q % 3]);
    }
    if (replay) {
        // light groups: dLoss/ds_m = (a_m - k P_m) / s_m, P_m = the group's total pmf
        gio->d_scale.assign(d.num_lights, 0.0);
        std::vector<double> P(d.num_lights, 0.0);
        for (size_t t = 0; t < sc.lights.size(); ++t) P[group_of(uint32_t(t))] += double(sc.light_pmf[t]);
        for (uint32_t m = 0; m < d.num_lights; ++m)
            if (d.lights[m].scale != 0) gio->d_scale[m] = (gio->a[m] - gio->k * P[m]) / double(d.lights[m].scale);
        // diffuse albedo RGB: d_rgb = J^T d_coef through the RGB-to-spectrum table
        gio->d_rgb.assign(size_t(d.num_materials) * 3, 0.0);
        for (uint32_t m = 0; m < d.num_materials; ++m) {
            if (d.materials[m].type != 0) continue;
            double J[3][3];
            sc.sp->table.jacobian(d.materials[m].rgb, J);
            for (int k = 0; k < 3; ++k)
                for (int q = 0; q < 3; ++q) gio->d_rgb[3 * size_t(m) + k] += J[q][k] * gio->d_coef[3 * size_t(m) + q];
        }
    }
    if (stats) {
        stats[0] = round;
        stats[1] = n_rays;
        stats[2] = n_shadow;
        stats[3] = n_done;
        stats[4] = uint64_t(std::chrono::duration_cast<std::chrono::microseconds>(
                                std::chrono::steady_clock::now() - t_start).count());
    }
}

// ------------------------------------------------------------ the megakernel integrators
// MegakernelIntegrator (base/megakernel_integrator.cu:14-42, :82-147): one thread
// per pixel renders that pixel's samples in sample order, one whole path at a
// time, and adds each to the film as it finishes.
constexpr int kIntMegakernelPath = 1, kIntAmbientOcclusion = 2, kIntSurfaceNormal = 3;

struct MkScene {
    const Scene &sc;
    uint32_t max_depth;
    float ao_scale;  // AmbientOcclusionIntegrator::create: 1 / illuminant.to_photometric(Y)
    Sigmoid sn[3];   // SurfaceNormalIntegrator::init: RGBAlbedoSpectrum of RGB 0.01 e_c
};
struct MkCount {
    uint64_t rays = 0, shadow = 0;
};

const float kInf = std::numeric_limits<float>::infinity();

inline bool mk_closest(const Scene &sc, V3 o, V3 d, uint32_t *prim, TriHit *h, MkCount &n) {
    float t, b[3];
    n.rays++;
    int rc = orc_trace_ray(sc.bvh, &o.x, &d.x, kInf, 0, prim, &t, b);
    if (rc < 0) throw std::runtime_error("oracle traversal stack overflow");
    *h = TriHit{b[0], b[1], b[2], t};
    return rc == 1;
}
inline bool mk_any(const Scene &sc, V3 o, V3 d, float t_max, MkCount &n) {
    uint32_t prim;
    n.shadow++;
    int rc = orc_trace_ray(sc.bvh, &o.x, &d.x, t_max, 1, &prim, nullptr, nullptr);
    if (rc < 0) throw std::runtime_error("oracle traversal stack overflow");
    return rc == 1;
}

// AmbientOcclusionIntegrator::li (integrators/ambient_occlusion.cu:25-58).
S4 ao_li(const MkScene &m, V3 ro, V3 rd, const Lambda &lam, Sampler &smp, MkCount &n) {
    uint32_t prim;
    TriHit h;
    if (!mk_closest(m.sc, ro, rd, &prim, &h, n)) return s4(0.0f);
    const Isect is = make_isect(m.sc, prim, h, rd);
    const V3 normal = face_forward(is.ns, neg(rd));
    float u0, u1;
    smp.get_2d(&u0, &u1);
    const V3 lw = sample_cosine_hemisphere(u0, u1);
    const float pdf = std::fabs(lw.z) / kPi;  // cosine_hemisphere_pdf (sampling.h:12-14)
    if (pdf == 0.0f) return s4(0.0f);
    V3 fx, fy;
    coordinate_system(normal, &fx, &fy);  // Frame::from_z (frame.h:24-29)
    const V3 wi = Frame{fx, fy, normal}.from_local(lw);
    if (mk_any(m.sc, offset_ray_origin(is.pi, is.n, wi), wi, kInf, n)) return s4(0.0f);
    return sample_pl(m.sc.sp->d65, lam) * (m.ao_scale * dot(normal, wi) / (kPi * pdf));
}

// SurfaceNormalIntegrator::li (integrators/surface_normal.cu:17-32); Vector3::softmax
// (vector3.h:134-137) with exp10 evaluated in double and rounded once to float.
S4 sn_li(const MkScene &m, V3 ro, V3 rd, const Lambda &lam, MkCount &n) {
    uint32_t prim;
    TriHit h;
    if (!mk_closest(m.sc, ro, rd, &prim, &h, n)) return s4(0.0f);
    const Isect is = make_isect(m.sc, prim, h, rd);
    const V3 v = normalize(face_forward(is.ns, neg(rd)));
    const float e[3] = {float(::exp10(double(v.x))), float(::exp10(double(v.y))), float(::exp10(double(v.z)))};
    const float s = e[0] + e[1] + e[2];
    const float c[3] = {e[0] / s, e[1] / s, e[2] / s};
    return (sample_sigmoid(m.sn[0], lam) * c[0] + sample_sigmoid(m.sn[1], lam) * c[1]) +
           sample_sigmoid(m.sn[2], lam) * c[2];
}

// MegakernelPathIntegrator::evaluate_li and sample_ld (integrators/megakernel_path.cu:35-213).
S4 mk_path_li(const MkScene &m, V3 ro, V3 rd, Lambda &lam, Sampler &smp, MkCount &n) {
    const Scene &sc = m.sc;
    S4 L = s4(0.0f), beta = s4(1.0f);
    bool specular = true;
    uint32_t depth = 0;
    float pdf_bsdf = NAN, eta_scale = 1;
    LightCtx prev{};
    for (;;) {
        uint32_t prim;
        TriHit h;
        if (!mk_closest(sc, ro, rd, &prim, &h, n)) {
            L = add_infinite(sc, L, beta, lam, depth, specular, pdf_bsdf);
            break;
        }
        const Tri &t = sc.tris[prim];
        const Isect is = make_isect(sc, prim, h, rd);
        if (t.light >= 0) {
            const S4 Le = light_L(sc, sc.lights[t.light], is.n, neg(rd), lam);
            if (positive(Le)) {
                if (depth == 0 || specular) {
                    L = L + beta * Le;
                } else {
                    const float pdf_light = sc.light_pmf[t.light] * emitter_pdf(sc, prim, prev, rd);
                    L = L + (beta * power_heuristic(pdf_bsdf, pdf_light)) * Le;
                }
            }
        }
        const Bsdf b = make_bsdf(sc, is, lam);  // before the depth test: may end the secondary wavelengths
        if (depth++ == m.max_depth) break;
        const uint32_t flags = bsdf_flags(b);
        if (non_specular(flags)) {
            LightCtx ctx{is.pi, is.n, is.ns};
            const bool refl = (flags & 1u) != 0, trans = (flags & 2u) != 0;
            if (refl && !trans) ctx.pi = make_p3fi(offset_ray_origin(is.pi, is.n, is.wo), {0, 0, 0});
            else if (trans && !refl) ctx.pi = make_p3fi(offset_ray_origin(is.pi, is.n, neg(is.wo)), {0, 0, 0});
            const float u = smp.get_1d();
            float ul0, ul1;
            smp.get_2d(&ul0, &ul1);
            if (!sc.lights.empty() && sc.lights[search_cdf(u, sc.select_cdf.data(), uint32_t(sc.lights.size()))].infinite) {
                const uint32_t li = search_cdf(u, sc.select_cdf.data(), uint32_t(sc.lights.size()));
                const V3 wi = sample_uniform_sphere(ul0, ul1);
                const S4 Le = light_L(sc, sc.lights[li], V3{0, 0, 0}, V3{0, 0, 0}, lam);
                if (positive(Le)) {
                    const S4 f = bsdf_f(b, is.wo, wi) * abs_dot(wi, is.ns);
                    if (positive(f)) {
                        const P3fi pl = make_p3fi(add(p3fi_mid(ctx.pi), scale(sc.inf_dist, wi)), {0, 0, 0});
                        V3 so[2], sd[2];
                        spawn_ray_to(is.pi, is.n, pl, V3{0, 0, 0}, &so[0], &sd[0]);
                        spawn_ray_to(pl, V3{0, 0, 0}, is.pi, is.n, &so[1], &sd[1]);
                        bool occ = false;
                        for (int r = 0; r < 2 && !occ; ++r) occ = mk_any(sc, so[r], sd[r], 0.6f, n);
                        if (!occ) {
                            const float pdf_light = sc.light_pmf[li] * kUniformSpherePdf;
                            const float wl = power_heuristic(pdf_light, bsdf_pdf(b, is.wo, wi));
                            L = L + beta * (((Le * wl) * f) / pdf_light);
                        }
                    }
                }
            } else if (!sc.lights.empty()) {
                const uint32_t li = search_cdf(u, sc.select_cdf.data(), uint32_t(sc.lights.size()));
                const Light &Lg = sc.lights[li];
                ShapeSample ss;
                const bool ok = emitter_sample(sc, Lg.tri, ctx, ul0, ul1, &ss);
                const V3 cpt = p3fi_mid(ctx.pi), sp = p3fi_mid(ss.pi);
                if (ok && !(ss.pdf == 0 || sqlen(sub(sp, cpt)) == 0)) {
                    const V3 wi = normalize(sub(sp, cpt));
                    const S4 Le = light_L(sc, Lg, ss.n, neg(wi), lam);
                    if (positive(Le) && ss.pdf != 0) {
                        const S4 f = bsdf_f(b, is.wo, wi) * abs_dot(wi, is.ns);
                        if (positive(f)) {
                            // IntegratorBase::unoccluded: p -> p_light, then p_light -> p
                            V3 so[2], sd[2];
                            spawn_ray_to(is.pi, is.n, ss.pi, ss.n, &so[0], &sd[0]);
                            spawn_ray_to(ss.pi, ss.n, is.pi, is.n, &so[1], &sd[1]);
                            bool occ = false;
                            for (int r = 0; r < 2 && !occ; ++r) occ = mk_any(sc, so[r], sd[r], 0.6f, n);
                            if (!occ) {
                                const float pdf_light = sc.light_pmf[li] * ss.pdf;
                                const float wl = power_heuristic(pdf_light, bsdf_pdf(b, is.wo, wi));
                                L = L + beta * (((Le * wl) * f) / pdf_light);
                            }
                        }
                    }
                }
            }
        }
        const V3 wo = neg(rd);
        const float uc = smp.get_1d();
        float u0, u1;
        smp.get_2d(&u0, &u1);
        BS bs;
        if (!bsdf_sample(b, wo, uc, u0, u1, &bs)) break;
        beta = beta * ((bs.f * abs_dot(bs.wi, is.ns)) / bs.pdf);
        pdf_bsdf = bs.pdf;
        specular = bs.specular;
        if (bs.transmission) eta_scale *= sqrf(bs.eta);
        prev = LightCtx{is.pi, is.n, is.ns};
        ro = offset_ray_origin(is.pi, is.n, bs.wi);
        rd = bs.wi;
        if (depth > 8) {  // Russian roulette: a sample is drawn only when beta's max is below 1
            const S4 rr = beta * eta_scale;
            if (max_comp(rr) < 1) {
                const float q = clampf(1 - max_comp(rr), 0, 0.95f);
                if (smp.get_1d() < q) break;
                beta = beta / (1 - q);
            }
        }
    }
    return L;
}

// path_L (nullable, [W*H*spp][4] by path id = sample * W * H + pixel): each
// sample's radiance before the sensor conversion.
void render_mk(const orc_scene &d, int integrator, int threads, float *film_rgb, float *aux, float *path_L,
               uint64_t *stats) {
    Scene sc;
    build_scene(sc, d);
    const uint32_t W = d.width, H = d.height;
    const uint32_t spd = uint32_t(std::sqrt(double(d.spp)));
    if (spd * spd != d.spp || d.spp == 0) throw std::runtime_error("spp must be a square");
    MkScene m{sc, d.max_depth, 1.0f / sc.sp->d65.inner(sc.sp->cie[1]), {}};
    for (int c = 0; c < 3; ++c) {
        float rgb[3] = {0, 0, 0};
        rgb[c] = 0.01f;
        m.sn[c] = sc.sp->table.lookup(rgb);
    }
    const uint64_t wh = uint64_t(W) * H;
    std::vector<float> rgb_sum(wh * 3, 0.f), w_sum(wh, 0.f);
    std::vector<MkCount> counts(wh);
    const auto t_start = std::chrono::steady_clock::now();
    parallel(wh, threads, [&](size_t px) {
        MkCount &n = counts[px];
        for (uint32_t s = 0; s < d.spp; ++s) {
            Sampler smp;
            smp.spd = spd;
            smp.start(uint32_t(px), s, 0);
            float u0, u1;
            smp.get_2d(&u0, &u1);
            float fx, fy, fw;
            sc.filter.sample(u0, u1, &fx, &fy, &fw);
            const float rx = (float(px % W) + fx) + 0.5f, ry = (float(px / W) + fy) + 0.5f;
            Lambda lam = sample_visible(smp.get_1d());
            const V3 ro = xf_point(sc.cam.render_from_camera.m, {0, 0, 0});
            const V3 rd = xf_vector(sc.cam.render_from_camera.m, normalize(xf_point(sc.cam.camera_from_raster.m, {rx, ry, 0})));
            S4 L;
            if (integrator == kIntAmbientOcclusion) L = ao_li(m, ro, rd, lam, smp, n);
            else if (integrator == kIntSurfaceNormal) L = sn_li(m, ro, rd, lam, n);
            else L = mk_path_li(m, ro, rd, lam, smp, n);
            if (path_L)
                for (int i = 0; i < kNS; ++i) path_L[4 * (uint64_t(s) * wh + px) + i] = L.v[i];
            float rgb[3];
            sensor_rgb(sc, L, lam, rgb);
            if (fw == 0) continue;
            for (int c = 0; c < 3; ++c) rgb_sum[3 * px + c] += fw * rgb[c];
            w_sum[px] += fw;
        }
    });
    for (uint64_t px = 0; px < wh; ++px) {
        float rgb[3] = {rgb_sum[3 * px], rgb_sum[3 * px + 1], rgb_sum[3 * px + 2]};
        if (w_sum[px] != 0)
            for (int c = 0; c < 3; ++c) rgb[c] /= w_sum[px];
        auto o = m3_apply(sc.out_from_sensor, rgb);
        for (int c = 0; c < 3; ++c) film_rgb[3 * px + c] = o[c];
        if (aux) {
            for (int c = 0; c < 3; ++c) aux[4 * px + c] = rgb_sum[3 * px + c];
            aux[4 * px + 3] = w_sum[px];
        }
    }
    if (stats) {
        MkCount tot;
        for (const MkCount &c : counts) tot.rays += c.rays, tot.shadow += c.shadow;
        stats[0] = 0;
        stats[1] = tot.rays;
        stats[2] = tot.shadow;
        stats[3] = wh * d.spp;
        stats[4] = uint64_t(std::chrono::duration_cast<std::chrono::microseconds>(
                                std::chrono::steady_clock::now() - t_start).count());
    }
}

}  // namespace
}  // namespace orc

extern "C" {

int orc_render(const orc_scene *scene, int threads, float *film_rgb, float *aux, uint64_t *stats) {
    if (!scene || !film_rgb || !scene->tri || !scene->tri_material || scene->pool_size == 0) return -1;
    try {
        orc::render(*scene, threads, film_rgb, aux, stats);
    } catch (const std::exception &) {
        return -2;
    }
    return 0;
}

int orc_render_megakernel(const orc_scene *scene, int integrator, int threads, float *film_rgb, float *aux,
                          float *path_L, uint64_t *stats) {
    if (!scene || !film_rgb || !scene->tri || !scene->tri_material) return -1;
    if (integrator < orc::kIntMegakernelPath || integrator > orc::kIntSurfaceNormal) return -1;
    try {
        orc::render_mk(*scene, integrator, threads, film_rgb, aux, path_L, stats);
    } catch (const std::exception &) {
        return -2;
    }
    return 0;
}

int orc_render_path_L(const orc_scene *scene, int threads, float *path_L, float *path_pdf) {
    if (!scene || !path_L || !scene->tri || !scene->tri_material || scene->pool_size == 0) return -1;
    try {
        std::vector<float> rgb(size_t(scene->width) * scene->height * 3);
        std::vector<orc::S4> L, pdf;
        orc::GradIO gio;
        gio.Lfin = &L;
        gio.pdffin = &pdf;
        orc::render(*scene, threads, rgb.data(), nullptr, nullptr, nullptr, nullptr, &gio);
        for (size_t i = 0; i < L.size(); ++i)
            for (int k = 0; k < orc::kNS; ++k) {
                path_L[4 * i + k] = L[i].v[k];
                if (path_pdf) path_pdf[4 * i + k] = pdf[i].v[k];
            }
    } catch (const std::exception &) {
        return -2;
    }
    return 0;
}

int orc_render_paths(const orc_scene *scene, int threads, uint32_t *path_start, uint32_t *path_finish,
                     uint64_t *stats) {
    if (!scene || !path_start || !path_finish || !scene->tri || scene->pool_size == 0) return -1;
    try {
        std::vector<float> rgb(size_t(scene->width) * scene->height * 3);
        orc::render(*scene, threads, rgb.data(), nullptr, stats, path_start, path_finish);
    } catch (const std::exception &) {
        return -2;
    }
    return 0;
}

int orc_render_grad(const orc_scene *scene, int threads, const float *target, float *film_rgb, double *loss,
                    double *d_rgb, double *d_scale, uint64_t *stats, int hold_selection) {
    if (!scene || !target || !film_rgb || !loss || !scene->tri || scene->pool_size == 0) return -1;
    try {
        const size_t npx = size_t(scene->width) * scene->height;
        std::vector<float> aux(npx * 4);
        std::vector<orc::S4> Lfin, pdffin;
        orc::GradIO fwd;
        fwd.Lfin = &Lfin;
        fwd.pdffin = &pdffin;
        orc::render(*scene, threads, film_rgb, aux.data(), stats, nullptr, nullptr, &fwd);
        // loss = sum (film - target)^2; A = d loss / d(rgb_sum / weight_sum) per pixel,
        // divided by the weight sum when the film divides by it
        std::vector<double> A(npx * 3, 0.0);
        double ls = 0;
        for (size_t px = 0; px < npx; ++px) {
            double gk[3];
            for (int k = 0; k < 3; ++k) {
                const double e = double(film_rgb[3 * px + k]) - double(target[3 * px + k]);
                ls += e * e;
                gk[k] = 2.0 * e;
            }
            const double W = aux[4 * px + 3];
            for (int c = 0; c < 3; ++c) {
                double v = 0;
                for (int k = 0; k < 3; ++k) v += gk[k] * fwd.M[3 * k + c];
                A[3 * px + c] = W != 0 ? v / W : v;
            }
        }
        *loss = ls;
        orc::GradIO rep;
        rep.hold_selection = hold_selection != 0;
        rep.Ltot = &Lfin;
        rep.pdf = &pdffin;
        rep.A = &A;
        std::vector<float> film2(npx * 3);
        orc::render(*scene, threads, film2.data(), nullptr, nullptr, nullptr, nullptr, &rep);
        if (d_rgb) std::copy(rep.d_rgb.begin(), rep.d_rgb.end(), d_rgb);
        if (d_scale) std::copy(rep.d_scale.begin(), rep.d_scale.end(), d_scale);
    } catch (const std::exception &) {
        return -2;
    }
    return 0;
}

int orc_rgb_to_sigmoid(const float rgb[3], float out3[3]) {
    orc::Sigmoid s = orc::spectra().table.lookup(rgb);
    out3[0] = s.c0;
    out3[1] = s.c1;
    out3[2] = s.c2;
    return 0;
}

int orc_rgb_to_sigmoid_jacobian(const float rgb[3], double J9[9]) {
    double J[3][3];
    orc::spectra().table.jacobian(rgb, J);
    for (int q = 0; q < 9; ++q) J9[q] = J[q / 3][q % 3];
    return 0;
}

int orc_srgb_matrices(float xyz_from_rgb9[9], float rgb_from_xyz9[9], float white_xy2[2]) {
    const orc::Spectra &s = orc::spectra();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            xyz_from_rgb9[3 * i + j] = s.xyz_from_rgb.m[i][j];
            rgb_from_xyz9[3 * i + j] = s.rgb_from_xyz.m[i][j];
        }
    white_xy2[0] = s.white_xy[0];
    white_xy2[1] = s.white_xy[1];
    return 0;
}

int orc_filter_sample(const float u[2], float out3[3]) {
    static orc::FilterSampler f;
    f.sample(u[0], u[1], &out3[0], &out3[1], &out3[2]);
    return 0;
}

float orc_spectrum_eval(int which, float lambda) {
    const orc::Spectra &s = orc::spectra();
    if (which >= 0 && which < 3) return s.cie[which](lambda);
    if (which == 3) return s.d65(lambda);
    static std::vector<orc::PLSpectrum> metals;
    static std::once_flag once;
    std::call_once(once, [] {
        const float *d[8] = {pbrtm_data::kAgEta, pbrtm_data::kAgK, pbrtm_data::kAlEta, pbrtm_data::kAlK,
                             pbrtm_data::kAuEta, pbrtm_data::kAuK, pbrtm_data::kCuEta, pbrtm_data::kCuK};
        const int n[8] = {pbrtm_data::kAgEtaLen, pbrtm_data::kAgKLen, pbrtm_data::kAlEtaLen, pbrtm_data::kAlKLen,
                          pbrtm_data::kAuEtaLen, pbrtm_data::kAuKLen, pbrtm_data::kCuEtaLen, pbrtm_data::kCuKLen};
        for (int i = 0; i < 8; ++i) metals.push_back(orc::pl_from_interleaved(d[i], n[i]));
    });
    if (which >= 4 && which < 12) return metals[which - 4](lambda);
    if (which == 12) {
        static const orc::PLSpectrum bk7 = orc::pl_from_interleaved(pbrtm_data::kGlassBK7Eta,
                                                                    pbrtm_data::kGlassBK7EtaLen);
        return bk7(lambda);
    }
    return NAN;
}

int orc_sample_visible(float u, float lam4[4], float pdf4[4]) {
    orc::Lambda w = orc::sample_visible(u);
    for (int i = 0; i < 4; ++i) lam4[i] = w.l[i], pdf4[i] = w.pdf[i];
    return 0;
}

int orc_sampler_sequence(uint32_t spd, uint32_t pixel, uint32_t sample, const int *kinds, int n, float *out) {
    orc::Sampler s;
    s.spd = spd;
    s.start(pixel, sample, 0);
    int k = 0;
    for (int i = 0; i < n; ++i) {
        if (kinds[i] == 1) {
            out[k++] = s.get_1d();
        } else {
            s.get_2d(&out[k], &out[k + 1]);
            k += 2;
        }
    }
    return k;
}

int orc_camera_ray(const orc_scene *sc, float px, float py, float out6[6]) {
    try {
        orc::Camera c = orc::make_camera(*sc);
        orc::V3 pc = orc::xf_point(c.camera_from_raster.m, {px, py, 0});
        orc::V3 o = orc::xf_point(c.render_from_camera.m, {0, 0, 0});
        orc::V3 d = orc::xf_vector(c.render_from_camera.m, orc::normalize(pc));
        orc::V3 ow = orc::xf_point(orc::xf_inverse(c.render_from_world).m, o);
        out6[0] = ow.x, out6[1] = ow.y, out6[2] = ow.z, out6[3] = d.x, out6[4] = d.y, out6[5] = d.z;
    } catch (const std::exception &) {
        return -1;
    }
    return 0;
}

int orc_spherical_triangle(const float v9[9], const float p[3], const float u[2], float b3[3], float *pdf) {
    orc::V3 v[3] = {{v9[0], v9[1], v9[2]}, {v9[3], v9[4], v9[5]}, {v9[6], v9[7], v9[8]}};
    orc::sample_spherical_triangle(b3, pdf, v, {p[0], p[1], p[2]}, u[0], u[1]);
    return 0;
}

int orc_invert_spherical_triangle(const float v9[9], const float p[3], const float w3[3], float u2[2]) {
    orc::V3 v[3] = {{v9[0], v9[1], v9[2]}, {v9[3], v9[4], v9[5]}, {v9[6], v9[7], v9[8]}};
    orc::invert_spherical_triangle(v, {p[0], p[1], p[2]}, {w3[0], w3[1], w3[2]}, &u2[0], &u2[1]);
    return 0;
}

float orc_fr_complex(float cos_i, float eta, float k) { return orc::fr_complex(cos_i, orc::Cx{eta, k}); }

// ---- shading functions one at a time, in the local frame (shading normal +z),
// for the mathematical pins of tests/test_oracle_bsdf.py.
static orc::Bsdf local_bsdf(const orc_bxdf &d) {
    orc::Bsdf b{};
    b.fr = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    b.type = d.type;
    b.tr = orc::TR{d.ax, d.ay};
    if (!b.tr.smooth()) b.tr = orc::TR{std::max(d.ax, 1e-4f), std::max(d.ay, 1e-4f)};
    for (int i = 0; i < orc::kNS; ++i) {
        b.r.v[i] = orc::clampf(d.r[i], 0.0f, 1.0f);
        b.eta.v[i] = d.eta4[i];
        b.k.v[i] = d.k4[i];
    }
    b.diffuse_set = orc::positive(b.r);
    b.eta_d = d.eta == 0 ? 1.0f : d.eta;
    return b;
}

uint32_t orc_bxdf_flags(const orc_bxdf *d) { return orc::bsdf_flags(local_bsdf(*d)); }

static int test_threads() { return int(std::max(1u, std::min(16u, std::thread::hardware_concurrency()))); }

int orc_bxdf_eval(const orc_bxdf *d, uint64_t n, const float *wo, const float *wi, float *f4, float *pdf) {
    const orc::Bsdf b = local_bsdf(*d);
    orc::parallel(n, test_threads(), [&](size_t i) {
        const orc::V3 o{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]}, w{wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]};
        const orc::S4 f = orc::bsdf_f(b, o, w);
        for (int k = 0; k < 4; ++k) f4[4 * i + k] = f.v[k];
        pdf[i] = orc::bsdf_pdf(b, o, w);
    });
    return 0;
}

int orc_bxdf_sample(const orc_bxdf *d, uint64_t n, const float *wo, const float *u3, float *f4, float *wi,
                    float *pdf, float *eta, uint32_t *flags) {
    const orc::Bsdf b = local_bsdf(*d);
    orc::parallel(n, test_threads(), [&](size_t i) {
        const orc::V3 o{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]};
        orc::BS bs{};
        const bool ok = orc::bsdf_sample(b, o, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2], &bs);
        flags[i] = ok ? (1u | (bs.specular ? 2u : 0u) | (bs.transmission ? 4u : 0u)) : 0u;
        for (int k = 0; k < 4; ++k) f4[4 * i + k] = ok ? bs.f.v[k] : 0.f;
        wi[3 * i] = ok ? bs.wi.x : 0.f;
        wi[3 * i + 1] = ok ? bs.wi.y : 0.f;
        wi[3 * i + 2] = ok ? bs.wi.z : 0.f;
        pdf[i] = ok ? bs.pdf : 0.f;
        eta[i] = ok ? bs.eta : 0.f;
    });
    return 0;
}

float orc_fr_dielectric(float cos_i, float eta) { return orc::fr_dielectric(cos_i, eta); }

int orc_refract(const float wi[3], const float n[3], float eta, float *etap, float wt[3]) {
    orc::V3 t{};
    if (!orc::refract({wi[0], wi[1], wi[2]}, {n[0], n[1], n[2]}, eta, etap, &t)) return 0;
    wt[0] = t.x, wt[1] = t.y, wt[2] = t.z;
    return 1;
}

int orc_tr_eval(float ax, float ay, uint64_t n, const float *w, const float *wm, float *out4) {
    const orc::TR tr{ax, ay};
    for (uint64_t i = 0; i < n; ++i) {
        const orc::V3 a{w[3 * i], w[3 * i + 1], w[3 * i + 2]}, m{wm[3 * i], wm[3 * i + 1], wm[3 * i + 2]};
        out4[4 * i] = tr.D(m);
        out4[4 * i + 1] = tr.lambda(a);
        out4[4 * i + 2] = tr.G1(a);
        out4[4 * i + 3] = tr.D(a, m);
    }
    return 0;
}

int orc_tr_sample(float ax, float ay, uint64_t n, const float *w, const float *u2, float *wm) {
    const orc::TR tr{ax, ay};
    for (uint64_t i = 0; i < n; ++i) {
        const orc::V3 m = tr.sample_wm({w[3 * i], w[3 * i + 1], w[3 * i + 2]}, u2[2 * i], u2[2 * i + 1]);
        wm[3 * i] = m.x, wm[3 * i + 1] = m.y, wm[3 * i + 2] = m.z;
    }
    return 0;
}

int orc_light_pmf(const orc_scene *d, float *pmf, int32_t *tri, uint32_t cap, uint32_t *count) {
    try {
        orc::Scene sc;
        orc::build_scene(sc, *d);
        *count = uint32_t(sc.lights.size());
        for (size_t i = 0; i < sc.lights.size() && i < cap; ++i) {
            pmf[i] = sc.light_pmf[i];
            tri[i] = sc.lights[i].infinite ? -1 : int32_t(sc.lights[i].tri);
        }
    } catch (const std::exception &) {
        return -1;
    }
    return 0;
}

// SRGBColorEncoding::from_linear (spectrum_util/color_encoding.h:7-60).
// ---- analytic shapes one at a time (tests): a shape with render_from_object m16
// and its inverse inv16 (row-major), reverse = ReverseOrientation.
static orc::quad::Quadric test_quadric(const orc_shape *sh, const float *m16, const float *inv16, int reverse) {
    using namespace orc;
    quad::Quadric q{};
    q.type = sh->type;
    std::memcpy(q.rfo.m, m16, sizeof(q.rfo.m));
    std::memcpy(q.rfo.inv, inv16, sizeof(q.rfo.inv));
    q.reverse = reverse != 0;
    const float phi_max = float(double(kPi) / 180.0 * double(clampf(sh->phimax, 0, 360)));
    const float r = sh->radius;
    q.radius = r;
    q.phi_max = phi_max;
    if (sh->type == 1) {
        q.z_min = clampf(std::min(sh->zmin, sh->zmax), -r, r);
        q.z_max = clampf(std::max(sh->zmin, sh->zmax), -r, r);
        q.theta_z_min = float(std::acos(double(clampf(q.z_min / r, -1, 1))));
        q.theta_z_max = float(std::acos(double(clampf(q.z_max / r, -1, 1))));
        q.flip = q.reverse;
    } else {
        q.height = sh->height;
        q.inner_radius = sh->innerradius;
        M4 m;
        std::memcpy(m.m, m16, sizeof(m.m));
        q.flip = q.reverse ^ xf_swaps_handedness(m);
    }
    return q;
}

/* Closest hit of each ray [n][8] with the shape: t (0 on a miss), the render-space hit
 * point and normal; returns the hit count. */
int orc_shape_intersect(const orc_shape *sh, const float *m16, const float *inv16, int reverse, const float *rays,
                        uint64_t n, float *t_out, float *p_out, float *n_out, uint8_t *hit_out) {
    using namespace orc;
    const quad::Quadric q = test_quadric(sh, m16, inv16, reverse);
    int hits = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const float *r = rays + 8 * i;
        quad::QuadricHit h;
        const bool hit = quad::quadric_intersect(q, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, r[3], &h, kPi);
        hit_out[i] = hit;
        t_out[i] = hit ? h.t : 0.f;
        if (hit) {
            const quad::QSurf g = quad::quadric_surface(q, h.p);
            const V3 pm = p3fi_mid(g.pi);
            for (int c = 0; c < 3; ++c) p_out[3 * i + c] = pm[c], n_out[3 * i + c] = g.n[c];
            ++hits;
        }
    }
    return hits;
}

/* Shape::sample(ctx, u) for reference points ctx [n][9] (p, n, ns) and u [n][2]:
 * the sampled point, its normal and solid-angle pdf; ok = 0 where sampling fails. */
int orc_shape_sample(const orc_shape *sh, const float *m16, const float *inv16, int reverse, const float *ctx,
                     const float *u, uint64_t n, float *p_out, float *n_out, float *pdf_out, uint8_t *ok_out) {
    using namespace orc;
    const quad::Quadric q = test_quadric(sh, m16, inv16, reverse);
    for (uint64_t i = 0; i < n; ++i) {
        const float *c = ctx + 9 * i;
        LightCtx lc{make_p3fi(V3{c[0], c[1], c[2]}, {0, 0, 0}), V3{c[3], c[4], c[5]}, V3{c[6], c[7], c[8]}};
        ShapeSample ss;
        const bool ok = q.type == 1 ? sphere_sample(q, lc, u[2 * i], u[2 * i + 1], &ss)
                                    : disk_sample(q, lc, u[2 * i], u[2 * i + 1], &ss);
        ok_out[i] = ok;
        const V3 pm = ok ? p3fi_mid(ss.pi) : V3{0, 0, 0};
        for (int k = 0; k < 3; ++k) p_out[3 * i + k] = pm[k], n_out[3 * i + k] = ok ? ss.n[k] : 0.f;
        pdf_out[i] = ok ? ss.pdf : 0.f;
    }
    return 0;
}

/* Shape::pdf(ctx, wi) for reference points ctx [n][9] and directions wi [n][3]. */
int orc_shape_pdf(const orc_shape *sh, const float *m16, const float *inv16, int reverse, const float *ctx,
                  const float *wi, uint64_t n, float *pdf_out) {
    using namespace orc;
    const quad::Quadric q = test_quadric(sh, m16, inv16, reverse);
    for (uint64_t i = 0; i < n; ++i) {
        const float *c = ctx + 9 * i;
        LightCtx lc{make_p3fi(V3{c[0], c[1], c[2]}, {0, 0, 0}), V3{c[3], c[4], c[5]}, V3{c[6], c[7], c[8]}};
        pdf_out[i] = quadric_pdf(q, lc, V3{wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]});
    }
    return 0;
}

/* The directed-rounding primitives of the interval arithmetic: out = [down, up] of
 * a * b (op 0), a / b (op 1), sqrt(a) (op 2), a + b (op 3). */
int orc_directed(int op, float a, float b, float out[2]) {
    using namespace orc::quad;
    switch (op) {
    case 0: out[0] = mul_rd(a, b), out[1] = mul_ru(a, b); return 0;
    case 1: out[0] = div_rd(a, b), out[1] = div_ru(a, b); return 0;
    case 2: out[0] = sqrt_rd(a), out[1] = sqrt_ru(a); return 0;
    case 3: out[0] = orc::add_round_down(a, b), out[1] = orc::add_round_up(a, b); return 0;
    }
    return -1;
}

int orc_srgb8(const float *rgb, uint64_t n, uint8_t *rgba) {
    auto to_srgb = [](float v) -> float {
        if (v <= 0.0031308f) return 12.92f * v;
        float sq = std::sqrt(std::max(0.0f, v));
        float p = std::fma(sq, std::fma(sq, std::fma(sq, std::fma(sq, std::fma(sq, -0.016202083165206348f,
                  0.7551545191665577f), 2.0041169284241644f), 0.7642611304733891f), 0.03453868659826638f),
                  -0.0016829072605308378f);
        float q = std::fma(sq, std::fma(sq, std::fma(sq, std::fma(sq, std::fma(sq, 1.f, 1.8970238036421054f),
                  0.6085338522168684f), 0.03467195408529984f), -0.00004375359692957097f), 4.178892964897981e-7f);
        return p / q * v;
    };
    for (uint64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            float v = rgb[3 * i + c];
            uint8_t o;
            if (v <= 0) o = 0;
            else if (v >= 1) o = 255;
            else if (std::isnan(v)) o = 0;
            else o = uint8_t(std::min(255.0, std::max(0.0, std::round(255.0 * double(to_srgb(v))))));
            rgba[4 * i + c] = o;
        }
        rgba[4 * i + 3] = 255;
    }
    return 0;
}

}  // extern "C"
