// Wavefront path-tracer kernels and the round loop that drives them.
//
// Reference: integrators/wavefront_path.cu (render :626-707, generate_new_path
// :265-308, control_logic :92-234, gpu_evaluate_material/sample_bsdf
// :310-431, sample_ld :574-624, write_frame_buffer :236-255). Differences in
// scheduling, none in results: control, material evaluation and new-path
// generation run fused in one kernel per round; the two shadow rays of a
// light sample are traced as a batch in the next round before the shading
// kernel adds the deferred contribution (the reference traces them inline;
// the contribution enters L at the same point of the sum); finished paths
// keep their film contribution per path id and the film kernel adds them per
// pixel in the reference's order (round, then sample).
#include "pathtrace.h"
#include "pbrtm_shapes.h"
#include "pt_device.h"
#include "wave_ops.h"

#include <algorithm>
#include <cstring>

namespace pbrtm {
namespace pt {

namespace {

constexpr uint32_t kBlock = 256;
// Ray ordering autotuner: rounds kTuneFirst .. kTuneFirst + 3 alternate sorted
// and unsorted queues (closest-hit and shadow rays separately); the cheaper
// time per ray over each pair decides the rest of the render. The film does not
// depend on the order.
constexpr uint32_t kTuneFirst = 2;
// Traversal occupancy autotuner: from round kOccFirst (after the ray order is
// settled), six rounds give the persistent traversal grids 4, 3, 2, 4, 3, 2
// quarters of the resident maximum; the least time per traced ray (closest-hit
// and shadow together) over each pair decides the rest of the render.
constexpr uint32_t kOccFirst = kTuneFirst + 5;
constexpr uint32_t kOccQuarters[3] = {4, 3, 2};
constexpr uint32_t kFlagSpecular = 1u << 16, kFlagAnyNonSpec = 1u << 17, kFlagNee = 1u << 18;
constexpr uint32_t kDepthMask = 0xFFFFu;
constexpr uint32_t kIdle = 0xFFFFFFFFu;

// counters[] layout (u32 words); 8..15 hold u64 accumulators of a batched render.
// A slot holds a live path exactly when its ray is active, so the closest-hit
// compaction's count (kCntQueue) is the round's live-path count.
constexpr int kCntNextId = 0, kCntQueue = 2, kCntSQueue = 3, kCntRound = 5, kCntDone = 6, kCntAccRays = 8;

struct State {
    uint32_t *path_id, *flags, *dim;
    uint64_t *rng;
    float4 *lam, *lam_pdf, *L, *beta, *nee_L, *scal, *prev;  // scal = (filter_w, pdf_bsdf, eta_scale, -)
    float4 *rays;                                             // [2 per slot]
    const uint32_t *prim;
    const float *thit, *bary;
    float4 *srays;                                            // [4 per slot]
    const uint32_t *sprim;                                    // [2 per slot]
    float4 *rec;                                              // per path id: sensor RGB, filter weight
    uint32_t *rec_round;     // finishing round; in a batch, rounds since the path started
    uint32_t *job, *sround;  // batch only: the slot's job and the round its path started
    float4 *lfin, *pfin;     // nullable: each path's final L and wavelength pdfs, by path id
    uint32_t *counters;
    uint32_t *err;
};

struct Params {
    uint64_t total;  // paths of the render, or of the whole batch
    uint32_t spd, spp, max_depth, round, pool;
    const JobDev *jobs;        // batch only
    const uint64_t *job_base;  // batch only: first path id of each job, then the batch total
    uint32_t njobs;
    GradDev gd;                // gradient replay only
};

// The job whose path-id range holds global path id `gid`.
__device__ __forceinline__ uint32_t find_job(const Params &P, uint64_t gid) {
    uint32_t lo = 0, hi = P.njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (P.job_base[mid] <= gid) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ f3 ld3(const float *p) { return {p[0], p[1], p[2]}; }

__device__ __forceinline__ f3 xf_point(const float *m, f3 p) {  // transform.h:254-265
    const float xp = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    const float yp = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    const float zp = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    const float wp = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    if (wp == 1.0f) return {xp, yp, zp};
    return {xp / wp, yp / wp, zp / wp};
}
__device__ __forceinline__ f3 xf_vector(const float *m, f3 v) {  // transform.h:303-307
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}

__device__ __forceinline__ float spec_eval(const SceneDev &S, uint32_t off, uint32_t n, float l) {
    return pl_eval(S.spec_lam + off, S.spec_val + off, n, l);
}

// pl_eval with the interval found by find_interval_guess (same result).
__device__ __forceinline__ float spec_eval_uni(const SceneDev &S, uint32_t off, uint32_t n, uint32_t a,
                                               float inv_step, float l) {
    const float *lam = S.spec_lam + off, *val = S.spec_val + off;
    if (n == 0 || l < lam[0] || l > lam[n - 1]) return 0;
    const size_t o = find_interval_guess(lam, n, a, inv_step, l);
    const float t = (l - lam[o]) / (lam[o + 1] - lam[o]);
    return lerpf(t, val[o], val[o + 1]);
}

// x, y, z of the CIE matching functions at l (one interval search when they share a grid).
__device__ __forceinline__ void cie_eval(const SceneDev &S, float l, float (&xyz)[3]) {
    if (!S.cie_shared) {
#pragma unroll
        for (int c = 0; c < 3; ++c) xyz[c] = spec_eval(S, S.cie_off[c], S.cie_n, l);
        return;
    }
    const uint32_t n = S.cie_n;
    const float *lam = S.spec_lam + S.cie_off[0];
    if (n == 0 || l < lam[0] || l > lam[n - 1]) {
        xyz[0] = xyz[1] = xyz[2] = 0.f;
        return;
    }
    const size_t o = find_interval_guess(lam, n, S.cie_uni, S.cie_inv_step, l);
    const float t = (l - lam[o]) / (lam[o + 1] - lam[o]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *val = S.spec_val + S.cie_off[c];
        xyz[c] = lerpf(t, val[o], val[o + 1]);
    }
}

// Piecewise-constant 1D sample (util/piecewise_constant_1d.h:61-86).
__device__ __forceinline__ float pc1d_sample(const float *func, const float *cdf, uint32_t n, float func_int,
                                             float mn, float mx, float u, float *pdf, int *off) {
    const int o = int(find_interval_le(cdf, n + 1, u));
    *off = o;
    float du = u - cdf[o];
    if (cdf[o + 1] - cdf[o] > 0) du /= cdf[o + 1] - cdf[o];
    *pdf = (func_int > 0) ? func[o] / func_int : 0;
    return lerpf((float(o) + du) / float(n), mn, mx);
}

// FilterSampler::sample (base/filter.h:66-73) over the tabulated Mitchell filter.
__device__ __forceinline__ void filter_sample(const SceneDev &S, float u0, float u1, float *px, float *py, float *w) {
    const uint32_t n = S.filt_n;
    const float r = S.filt_radius;
    float pdf0, pdf1;
    int o0, o1;
    const float d1 = pc1d_sample(S.filt_marg_func, S.filt_marg_cdf, n, S.filt_marg_int, -r, r, u1, &pdf1, &o1);
    const float d0 = pc1d_sample(S.filt_cond_func + size_t(o1) * n, S.filt_cond_cdf + size_t(o1) * (n + 1), n,
                                 S.filt_cond_int[o1], -r, r, u0, &pdf0, &o0);
    *px = d0;
    *py = d1;
    *w = S.filt_f[o0 + n * o1] / (pdf0 * pdf1);
}

// distribution_1d.h:9-42
__device__ __forceinline__ uint32_t search_cdf(float u, const float *cdf, uint32_t len) {
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
        const uint32_t mid = (start + end) / 2;
        if (u >= cdf[mid]) start = mid + 1;
        else end = mid + 1;
    }
}

struct Tri {
    f3 p0, p1, p2;
    bool flip;
};
__device__ __forceinline__ Tri load_tri(const SceneDev &S, uint32_t i) {
    const float *t = S.tri + size_t(i) * 9;
    return {ld3(t), ld3(t + 3), ld3(t + 6), (S.tri_info[i].y >> 31) != 0};
}
__device__ __forceinline__ float tri_area(const Tri &t) {
    return len3(cross3(sub3(t.p1, t.p0), sub3(t.p2, t.p0))) * 0.5f;
}
__device__ __forceinline__ f3 tri_normal(const Tri &t) {  // shapes/triangle.cu:391-395
    f3 n = norm3(cross3(sub3(t.p0, t.p2), sub3(t.p1, t.p2)));
    return t.flip ? neg3(n) : n;
}
__device__ __forceinline__ f3 light_normal(const Tri &t) {  // shapes/triangle.cu:185-194
    f3 n = norm3(cross3(sub3(t.p1, t.p0), sub3(t.p2, t.p0)));
    return t.flip ? neg3(n) : n;
}

// DiffuseAreaLight::l (lights/diffuse_area_light.cu:37-46) with RGBIlluminantSpectrum.
__device__ __forceinline__ S4 light_L(const SceneDev &S, const DLight &L, f3 n, f3 w, const S4 &lam) {
    if (!L.two_sided && dot3(n, w) < 0.0f) return s4(0.0f);
    S4 s, d;
    for (int i = 0; i < kNS; ++i) {
        s.v[i] = L.rgb_scale * sigmoid_eval(L.c0, L.c1, L.c2, lam.v[i]);
        d.v[i] = spec_eval_uni(S, S.d65_off, S.d65_n, S.d65_uni, S.d65_inv_step, lam.v[i]);
    }
    return s4scale(s4mul(d, s), L.scale);
}

struct ShapeSample {
    P3i pi;
    f3 n;
    float pdf;
};
__device__ __forceinline__ P3i sample_point(const Tri &t, const float b[3]) {
    const f3 pabs = add3(add3(abs3(mul3(b[0], t.p0)), abs3(mul3(b[1], t.p1))), abs3(mul3(1 - b[0] - b[1], t.p2)));
    const f3 perr = mul3(gamma_n(6), pabs);
    const f3 p = add3(add3(mul3(b[0], t.p0), mul3(b[1], t.p1)), mul3(b[2], t.p2));
    return p3i_make(p, perr);
}
// Triangle::sample(ctx, u) (shapes/triangle.cu:126-215) and its area-sampling branch (:78-124).
__device__ __forceinline__ bool tri_sample(const Tri &t, const P3i &cpi, f3 cns, float u0, float u1, ShapeSample *ss) {
    const f3 cp = p3i_mid(cpi);
    const f3 v[3] = {t.p0, t.p1, t.p2};
    const float sa = spherical_triangle_area(norm3(sub3(t.p0, cp)), norm3(sub3(t.p1, cp)), norm3(sub3(t.p2, cp)));
    if (sa < kMinSolidAngle || sa > kMaxSolidAngle) {
        float b[3];
        if (u0 < u1) {
            b[0] = u0 / 2;
            b[1] = u1 - b[0];
        } else {
            b[1] = u1 / 2;
            b[0] = u0 - b[1];
        }
        b[2] = 1 - b[0] - b[1];
        ss->pi = sample_point(t, b);
        ss->n = light_normal(t);
        ss->pdf = float(1.0) / tri_area(t);
        const f3 pp = p3i_mid(ss->pi);
        f3 wi = sub3(pp, cp);
        if (sqlen3(wi) == 0) return false;
        wi = norm3(wi);
        ss->pdf /= absdot3(ss->n, neg3(wi)) / sqlen3(sub3(cp, pp));
        if (isinf(ss->pdf)) return false;
        return true;
    }
    float pdf = 1.0f;
    if (!zero3(cns)) {
        const f3 w0 = norm3(sub3(t.p0, cp)), w1 = norm3(sub3(t.p1, cp)), w2 = norm3(sub3(t.p2, cp));
        const float w[4] = {smax(0.01f, absdot3(cns, w1)), smax(0.01f, absdot3(cns, w1)), smax(0.01f, absdot3(cns, w0)),
                            smax(0.01f, absdot3(cns, w2))};
        const float y = sample_linear(u1, w[0] + w[1], w[2] + w[3]);
        const float x = sample_linear(u0, lerpf(y, w[0], w[2]), lerpf(y, w[1], w[3]));
        u0 = x;
        u1 = y;
        pdf = bilinear_pdf(u0, u1, w);
    }
    float b[3], tpdf;
    sample_spherical_triangle(b, &tpdf, v, cp, u0, u1);
    if (tpdf == 0.0f) return false;
    pdf *= tpdf;
    ss->pi = sample_point(t, b);
    ss->n = light_normal(t);
    ss->pdf = pdf;
    return true;
}
// Triangle::pdf(ctx, wi) (shapes/triangle.cu:20-76).
__device__ __forceinline__ float tri_pdf(const Tri &t, const P3i &cpi, f3 cn, f3 cns, f3 wi) {
    const f3 cp = p3i_mid(cpi);
    const f3 v[3] = {t.p0, t.p1, t.p2};
    const float sa = spherical_triangle_area(norm3(sub3(t.p0, cp)), norm3(sub3(t.p1, cp)), norm3(sub3(t.p2, cp)));
    if (sa < kMinSolidAngle || sa > kMaxSolidAngle) {
        const f3 o = ctx_offset_origin(cpi, cn, wi);
        TriHit h;
        if (!tri_hit(o, wi, __builtin_inff(), t.p0, t.p1, t.p2, &h)) return 0;
        const HitGeom g = hit_geom(h, t.p0, t.p1, t.p2);
        const f3 n = t.flip ? neg3(g.n) : g.n;
        const f3 ip = {ival_mid(g.x), ival_mid(g.y), ival_mid(g.z)};
        float pdf = (1 / tri_area(t)) / (absdot3(n, neg3(wi)) / sqlen3(sub3(cp, ip)));
        if (isinf(pdf)) pdf = 0;
        return pdf;
    }
    float pdf = 1 / sa;
    if (!zero3(cns)) {
        float u0, u1;
        invert_spherical_triangle(v, cp, wi, &u0, &u1);
        const f3 w0 = norm3(sub3(t.p0, cp)), w1 = norm3(sub3(t.p1, cp)), w2 = norm3(sub3(t.p2, cp));
        const float w[4] = {smax(0.01f, absdot3(cns, w1)), smax(0.01f, absdot3(cns, w1)), smax(0.01f, absdot3(cns, w0)),
                            smax(0.01f, absdot3(cns, w2))};
        pdf *= bilinear_pdf(u0, u1, w);
    }
    return pdf;
}

// Ray::spawn_ray_to (base/ray.h:45-51): origin pf and unnormalised direction pt - pf.
__device__ __forceinline__ void spawn_ray_to(const P3i &from, f3 n_from, const P3i &to, f3 n_to, f3 *o, f3 *d) {
    const f3 pf = offset_ray_origin(from, n_from, sub3(p3i_mid(to), p3i_mid(from)));
    const f3 pt = offset_ray_origin(to, n_to, sub3(pf, p3i_mid(to)));
    *o = pf;
    *d = sub3(pt, pf);
}

// The surface interaction at a closest hit: a triangle (shapes/triangle.cu:330-470,
// default uv, no shading normals), or an analytic primitive of S.shapes (spheres,
// disks, instanced triangles), whose hit point or base barycentrics the traversal
// left in the barycentric slots.
__device__ __forceinline__ Surf surface_at(const SceneDev &S, uint32_t prim, const TriHit &th) {
    if (S.prim_shape) {
        const uint32_t k = S.prim_shape[prim];
        if (k) return shape_surf(S.shapes[k - 1], f3{th.b0, th.b1, th.b2});
    }
    const Tri tri = load_tri(S, prim);
    const f3 dp02 = sub3(tri.p0, tri.p2), dp12 = sub3(tri.p1, tri.p2);
    f3 dpdu = {dop(-1.f, dp02.x, -1.f, dp12.x), dop(-1.f, dp02.y, -1.f, dp12.y), dop(-1.f, dp02.z, -1.f, dp12.z)};
    f3 dpdv = {dop(-1.f, dp12.x, 0.f, dp02.x), dop(-1.f, dp12.y, 0.f, dp02.y), dop(-1.f, dp12.z, 0.f, dp02.z)};
    dpdu = mul3(1.0f, dpdu);
    dpdv = mul3(1.0f, dpdv);
    if (sqlen3(cross3(dpdu, dpdv)) == 0) basis(norm3(cross3(sub3(tri.p2, tri.p0), sub3(tri.p1, tri.p0))), &dpdu, &dpdv);
    const HitGeom g = hit_geom(th, tri.p0, tri.p1, tri.p2);
    const f3 n = tri.flip ? neg3(g.n) : g.n;
    return Surf{P3i{g.x, g.y, g.z}, n, n, dpdu};
}

// DiffuseAreaLight::pdf_li (lights/diffuse_area_light.cu:69-74) through the emitter's shape.
__device__ __forceinline__ float prim_light_pdf(const SceneDev &S, uint32_t prim, const P3i &cpi, f3 cn, f3 cns,
                                                f3 wi) {
    if (S.prim_shape) {
        const uint32_t k = S.prim_shape[prim];
        if (k) return shape_light_pdf(S.shapes[k - 1], cpi, cn, wi);
    }
    return tri_pdf(load_tri(S, prim), cpi, cn, cns, wi);
}

// Shape::sample(ctx, u) of the emitter's shape (DiffuseAreaLight::sample_li,
// lights/diffuse_area_light.cu:48-67).
__device__ __forceinline__ bool prim_light_sample(const SceneDev &S, uint32_t prim, const P3i &cpi, f3 cn, f3 cns,
                                                  float u0, float u1, ShapeSample *ss) {
    if (S.prim_shape) {
        const uint32_t k = S.prim_shape[prim];
        if (k) {
            QSample q;
            if (!shape_light_sample(S.shapes[k - 1], cpi, cn, u0, u1, &q)) return false;
            ss->pi = q.pi;
            ss->n = q.n;
            ss->pdf = q.pdf;
            return true;
        }
    }
    return tri_sample(load_tri(S, prim), cpi, cns, u0, u1, ss);
}

__device__ __forceinline__ void store_ray(float4 *r, f3 o, float tmax, f3 d) {
    r[0] = make_float4(o.x, o.y, o.z, tmax);
    r[1] = make_float4(d.x, d.y, d.z, 0.f);
}

// PixelSensor::to_sensor_rgb (films/pixel_sensor.h:33-40), imaging ratio 1.
__device__ __forceinline__ float3 sensor_rgb(const SceneDev &S, const S4 &L, const S4 &lam, const S4 &pdf) {
    const S4 s = s4safediv(L, pdf);
    float xyz[kNS][3];
#pragma unroll
    for (int i = 0; i < kNS; ++i) cie_eval(S, lam.v[i], xyz[i]);
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S4 x;
#pragma unroll
        for (int i = 0; i < kNS; ++i) x.v[i] = xyz[i][k];
        c[k] = 1.0f * s4avg(s4mul(x, s));
    }
    return make_float3(c[0], c[1], c[2]);
}

// generate_new_path (wavefront_path.cu:265-308) for path `id` of the render
// (global id `gid` in a batch) in slot `i`.
template <bool GRAD = false>
__device__ __forceinline__ void start_path(const SceneDev &S, const State &st, uint32_t spd, uint32_t i, uint64_t id, uint64_t gid,
                           const float4 *g_lfin = nullptr, const float4 *g_pfin = nullptr, const float *g_adj = nullptr,
                           float4 *g_g = nullptr, float4 *g_ltot = nullptr) {
    const uint64_t wh = uint64_t(S.width) * S.height;
    const uint32_t pixel = uint32_t(id % wh), sample = uint32_t(id / wh);
    Sampler smp{pcg_seq(hash_u32(pixel)), pixel, sample, 0, spd};
    pcg_advance(smp.rng, uint64_t(sample) * 65536ull);
    float u0, u1;
    smp.get_2d(&u0, &u1);
    float fx, fy, fw;
    filter_sample(S, u0, u1, &fx, &fy, &fw);
    const float px = (float(pixel % S.width) + fx) + 0.5f;
    const float py = (float(pixel / S.width) + fy) + 0.5f;
    S4 lam, pdf;
    sample_wavelengths(smp.get_1d(), &lam, &pdf);
    const f3 pc = xf_point(S.cam_from_raster, {px, py, 0.f});
    const f3 dc = norm3(pc);
    const f3 rd = xf_vector(S.render_from_camera, dc);
    st.path_id[i] = uint32_t(gid);
    st.flags[i] = kFlagSpecular;
    st.rng[i] = smp.rng.state;
    st.dim[i] = smp.dim;
    st.lam[i] = s4f(lam);
    st.lam_pdf[i] = s4f(pdf);
    st.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    st.beta[i] = make_float4(1.f, 1.f, 1.f, 1.f);
    st.scal[i] = make_float4(fw, __builtin_nanf(""), 1.0f, 0.f);
    store_ray(st.rays + 2 * size_t(i), ld3(S.cam_origin), __builtin_inff(), rd);
    if (GRAD) {
        // dLoss/dL_i = w / 4 * sum_c A_c x_c(lambda_i) / pdf_i, pdf_i the path's final one
        // (PixelSensor::to_sensor_rgb, RGBFilm::add_sample and get_pixel_rgb)
        const S4 pf = f4s(g_pfin[id]);
        const float *A = g_adj + 3 * size_t(pixel);
        S4 g;
#pragma unroll
        for (int k = 0; k < kNS; ++k) {
            float v = 0.f;
            if (pf.v[k] != 0.f) {
                float xyz[3];
                cie_eval(S, lam.v[k], xyz);
#pragma unroll
                for (int c = 0; c < 3; ++c) v += A[c] * xyz[c] / pf.v[k];
            }
            g.v[k] = fw * v * (1.0f / kNS);
        }
        g_g[i] = s4f(g);
        g_ltot[i] = g_lfin[id];
    }
}

__device__ __forceinline__ void idle_slot(const State &st, uint32_t i) {
    st.path_id[i] = kIdle;
    st.rays[2 * size_t(i)] = make_float4(0.f, 0.f, 0.f, 0.f);
}

template <bool GRAD>
__global__ __launch_bounds__(kBlock) void init_pool_kernel(SceneDev S, State st, Params P) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.pool) return;
    st.srays[4 * size_t(i)] = make_float4(0.f, 0.f, 0.f, 0.f);
    st.srays[4 * size_t(i) + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (uint64_t(i) < P.total)
        start_path<GRAD>(S, st, P.spd, i, i, i, P.gd.lfin, P.gd.pfin, P.gd.adj, P.gd.g, P.gd.ltot);
    else idle_slot(st, i);
}

__global__ __launch_bounds__(kBlock) void init_pool_multi_kernel(State st, Params P) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.pool) return;
    st.srays[4 * size_t(i)] = make_float4(0.f, 0.f, 0.f, 0.f);
    st.srays[4 * size_t(i) + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (uint64_t(i) < P.total) {
        const uint32_t j = find_job(P, i);
        const JobDev &J = P.jobs[j];
        start_path(J.S, st, J.spd, i, i - J.base, i);
        st.job[i] = j;
        st.sround[i] = 0;
    } else {
        idle_slot(st, i);
    }
}

// One round for slot i: deferred NEE, control_logic, film record, new path,
// material evaluation with sample_ld and sample_f. pid is the slot's path id
// (global in a batch), lid its id within its render, S its scene; every lane
// of the wave calls this (the new-path ballot is wave-wide).
// Gradient replay accumulators: s_acc[group] gathers the scale-linear part of
// dLoss/d(emitter scale) and s_acc[ngroups] the light-pmf part (shared memory,
// flushed per block).
__device__ __forceinline__ float dot4(const S4 &a, const S4 &b) {
    float v = 0.f;
#pragma unroll
    for (int k = 0; k < kNS; ++k) v += a.v[k] * b.v[k];
    return v;
}

// MODE is the integrator: kIntPath is the wavefront path integrator; the others
// run the megakernel integrators (base/megakernel_integrator.cu:82-147) on the
// same pool, one round per path vertex.
template <bool MULTI, bool GRAD, uint32_t MODE = kIntPath>
__device__ __forceinline__ void shade_impl(const SceneDev &S, const State &st, const Params &P, uint32_t i,
                                           uint32_t pid, uint32_t lid, uint32_t spd, uint32_t max_depth,
                                           uint32_t round, float4 *erec, float4 *grec) {
    const bool live = pid != kIdle;
    uint32_t fl = live ? st.flags[i] : 0u;
    S4 L = live ? f4s(st.L[i]) : s4(0.f), beta = live ? f4s(st.beta[i]) : s4(0.f);
    S4 gadj = s4(0.f);
    if (GRAD && live) gadj = f4s(P.gd.g[i]);
    // 1. the previous round's light sample, if both shadow rays missed
    if (live && (fl & kFlagNee)) {
        if (st.sprim[2 * size_t(i)] == kNoHit && st.sprim[2 * size_t(i) + 1] == kNoHit) {
            L = s4add(L, f4s(st.nee_L[i]));
            if (GRAD) {
                // C = beta f cos Le w(pmf p) / (pmf p) with w = g^2 / (g^2 + b^2), g = pmf p:
                // dC/dpmf = C (1 - 2w) / pmf (the MIS weight's pmf and the light choice's
                // 1/pmf both attached), so a += G 2 (1 - w) and k += G (1 - 2w)
                const float4 q = P.gd.nee[i];
                erec[0] = make_float4(q.x * 2.f * (1.f - q.y), q.x * (1.f - 2.f * q.y), q.z, 0.f);
            }
        }
        fl &= ~kFlagNee;
    }
    const uint32_t prim = live ? st.prim[i] : kNoHit;
    const uint32_t depth = fl & kDepthMask;
    bool terminate = live && (prim == kNoHit || depth > max_depth || !s4pos(beta));
    Sampler smp{};
    S4 lam = s4(0.f), lpdf = s4(0.f);
    float4 scal = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        smp.pixel = lid % (S.width * S.height);
        smp.sample = lid / (S.width * S.height);
        smp.spd = spd;
        smp.rng.state = st.rng[i];
        smp.rng.inc = sampler_inc(smp.pixel);
        smp.dim = st.dim[i];
        lam = f4s(st.lam[i]);
        lpdf = f4s(st.lam_pdf[i]);
        scal = st.scal[i];
    }
    // Russian roulette (wavefront_path.cu:121-138), only past 8 bounces. The
    // wavefront draws its sample always; the megakernel path integrator only
    // when beta's largest value is below 1 (megakernel_path.cu:142-153).
    if (live && !terminate && depth > 8) {
        const float u = MODE == kIntMegakernelPath ? 0.f : smp.get_1d();
        const S4 rr = s4scale(beta, scal.z);
        if (s4max(rr) < 1) {
            const float q = clampf(1 - s4max(rr), 0, 0.95f);
            if ((MODE == kIntMegakernelPath ? smp.get_1d() : u) < q) {
                beta = s4(0.0f);
                terminate = true;
            } else {
                beta = s4div(beta, 1 - q);
            }
        }
    }
    // Emission with MIS at this vertex (wavefront_path.cu:179-196,
    // megakernel_path.cu:81-99).
    auto add_emission = [&](uint32_t prim, uint32_t light, f3 n, f3 rd) {
        const DLight Ld = S.lights[light - 1];
        const S4 Le = light_L(S, Ld, n, neg3(rd), lam);
        if (!s4pos(Le)) return;
        if (depth == 0 || (fl & kFlagSpecular)) {
            L = s4add(L, s4mul(beta, Le));
            if (GRAD) erec[1] = make_float4(dot4(gadj, s4mul(beta, Le)), 0.f, __uint_as_float(Ld.group), 0.f);
        } else {
            const float4 q0 = st.prev[3 * size_t(i)], q1 = st.prev[3 * size_t(i) + 1], q2 = st.prev[3 * size_t(i) + 2];
            const P3i ppi{{q0.x, q0.y}, {q0.z, q0.w}, {q1.x, q1.y}};
            const f3 pn = {q1.z, q1.w, q2.x}, pns = {q2.y, q2.z, q2.w};
            const float pdf_light = S.light_pmf[light - 1] * prim_light_pdf(S, prim, ppi, pn, pns, rd);
            const float wl = power_heuristic(scal.y, pdf_light);
            L = s4add(L, s4mul(s4scale(beta, wl), Le));
            if (GRAD) {  // MIS weight attached: dC/dpmf = -2 C (1 - w) / pmf
                const double G = dot4(gadj, s4mul(s4scale(beta, wl), Le));
                erec[1] = make_float4(float(G * (2.0 * double(wl) - 1.0)), float(-2.0 * G * (1.0 - double(wl))),
                                      __uint_as_float(Ld.group), 0.f);
            }
        }
    };
    // The megakernel path integrator ends a path at depth max_depth after the
    // vertex's emission and get_bsdf (megakernel_path.cu:81-111): no light
    // sample, no BSDF sample. get_bsdf of a dielectric with a spectral eta still
    // ends the secondary wavelengths (dielectric_material.cu:47-50).
    if (MODE == kIntMegakernelPath && live && !terminate && depth == max_depth) {
        const uint2 info = S.tri_info[prim];
        const TriHit th{st.bary[3 * size_t(i)], st.bary[3 * size_t(i) + 1], st.bary[3 * size_t(i) + 2], st.thit[i]};
        const float4 r1 = st.rays[2 * size_t(i) + 1];
        const uint32_t light = info.y & 0x7FFFFFFFu;
        if (light) add_emission(prim, light, surface_at(S, prim, th).n, f3{r1.x, r1.y, r1.z});
        const DMat m = S.mats[info.x];
        if (m.type == kDielectric && m.eta_n && !(lpdf.v[1] == 0 && lpdf.v[2] == 0 && lpdf.v[3] == 0)) {
            lpdf.v[1] = lpdf.v[2] = lpdf.v[3] = 0;
            lpdf.v[0] /= kNS;
        }
        terminate = true;
    }
    // 2. finished paths: film record, then a new path in this slot. The new path
    // ids are reserved once per workgroup (every thread of it gets here): the ids
    // handed out in a round are one contiguous run whatever the order of the
    // reservations, and one shared counter serves them one at a time.
    const uint64_t lanes = __ballot(terminate);
    __shared__ uint32_t s_new[kBlock / kWave];
    __shared__ uint32_t s_new_base;
    const uint32_t wv = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
    if (lane_id() == 0) s_new[wv] = uint32_t(__popcll(lanes));
    __syncthreads();
    uint32_t before = 0, total_new = 0;
    for (uint32_t k = 0; k < nwaves; ++k) {
        const uint32_t t = s_new[k];
        before += k < wv ? t : 0u;
        total_new += t;
    }
    if (threadIdx.x == 0) s_new_base = total_new ? atomicAdd(&st.counters[kCntNextId], total_new) : 0u;
    __syncthreads();
    const uint32_t base = s_new_base + before;
    if (terminate) {
        // Infinite lights (wavefront_path.cu:140-160): every path that ends with a
        // positive beta, escaped or not; the megakernel path integrator only on an
        // escape (megakernel_path.cu:58-75). pdf_li allowing an incomplete pdf is 0.
        if (MODE <= kIntMegakernelPath && S.num_inf && (MODE == kIntPath ? s4pos(beta) : prim == kNoHit)) {
            for (uint32_t k = 0; k < S.num_inf; ++k) {
                const uint32_t li = S.inf_idx[k];
                const S4 Le = light_L(S, S.lights[li], f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}, lam);
                if (depth == 0 || (fl & kFlagSpecular)) L = s4add(L, s4mul(beta, Le));
                else L = s4add(L, s4mul(s4scale(beta, power_heuristic(scal.y, S.light_pmf[li] * 0.0f)), Le));
            }
        }
        const float3 rgb = sensor_rgb(S, L, lam, lpdf);
        st.rec[pid] = make_float4(rgb.x, rgb.y, rgb.z, scal.x);
        // the megakernel integrators add each pixel's samples in sample order
        st.rec_round[pid] = MODE != kIntPath ? 0u : (MULTI ? round - st.sround[i] : round);
        if (st.lfin) {
            st.lfin[pid] = s4f(L);
            st.pfin[pid] = s4f(lpdf);
        }
        const uint64_t nid = uint64_t(base) + __popcll(lanes & lanemask_lt());
        if (nid < P.total) {
            if (MULTI) {
                const uint32_t j = find_job(P, nid);
                const JobDev &J = P.jobs[j];
                start_path(J.S, st, J.spd, i, nid - J.base, nid);
                st.job[i] = j;
                st.sround[i] = round + 1;
            } else {
                start_path<GRAD>(S, st, spd, i, nid, nid, P.gd.lfin, P.gd.pfin, P.gd.adj, P.gd.g, P.gd.ltot);
            }
        } else {
            idle_slot(st, i);
        }
        st.srays[4 * size_t(i)].w = 0.f;
        st.srays[4 * size_t(i) + 2].w = 0.f;
        return;
    }
    if (!live) return;
    const float4 r1 = st.rays[2 * size_t(i) + 1];
    const f3 rd = {r1.x, r1.y, r1.z};
    const uint2 info = S.tri_info[prim];
    const TriHit th{st.bary[3 * size_t(i)], st.bary[3 * size_t(i) + 1], st.bary[3 * size_t(i) + 2], st.thit[i]};
    const Surf sf = surface_at(S, prim, th);
    const P3i ipi = sf.pi;
    const f3 n = sf.n, ns = sf.ns, dpdu = sf.dpdu;
    if (MODE == kIntAmbientOcclusion || MODE == kIntSurfaceNormal) {
        // One vertex, no emission; the path ends next round (its ray is retired).
        const f3 nf = dot3(ns, neg3(rd)) < 0.0f ? neg3(ns) : ns;  // Vector3::face_forward
        if (MODE == kIntSurfaceNormal) {
            // SurfaceNormalIntegrator::li (surface_normal.cu:17-32): softmax of the
            // normal (exp10 in double, rounded once) weighting three albedo spectra
            const f3 v = norm3(nf);
            const float e0 = float(::exp10(double(v.x))), e1 = float(::exp10(double(v.y))),
                        e2 = float(::exp10(double(v.z)));
            const float es = e0 + e1 + e2;
            const float c0 = e0 / es, c1 = e1 / es, c2 = e2 / es;
#pragma unroll
            for (int k = 0; k < kNS; ++k) {
                const float l = lam.v[k];
                L.v[k] = (sigmoid_eval(S.sn_c[0], S.sn_c[1], S.sn_c[2], l) * c0 +
                          sigmoid_eval(S.sn_c[3], S.sn_c[4], S.sn_c[5], l) * c1) +
                         sigmoid_eval(S.sn_c[6], S.sn_c[7], S.sn_c[8], l) * c2;
            }
        } else {
            // AmbientOcclusionIntegrator::li (ambient_occlusion.cu:25-58): a cosine
            // sample around the face-forwarded normal, its any-hit ray traced next
            // round with t_max = infinity
            float u0, u1;
            smp.get_2d(&u0, &u1);
            const f3 lw = cosine_hemisphere(u0, u1);
            const float pdf = fabsf(lw.z) / kPi;
            bool ao = false;
            if (pdf != 0.0f) {
                f3 bx, by;
                basis(nf, &bx, &by);
                const f3 wi = add3(add3(mul3(lw.x, bx), mul3(lw.y, by)), mul3(lw.z, nf));
                const float k = S.ao_scale * dot3(nf, wi) / (kPi * pdf);
                S4 d;
#pragma unroll
                for (int q = 0; q < kNS; ++q)
                    d.v[q] = spec_eval_uni(S, S.d65_off, S.d65_n, S.d65_uni, S.d65_inv_step, lam.v[q]) * k;
                st.nee_L[i] = s4f(d);
                store_ray(st.srays + 4 * size_t(i), offset_ray_origin(ipi, n, wi), __builtin_inff(), wi);
                st.srays[4 * size_t(i) + 2].w = 0.f;
                ao = true;
            }
            if (ao) fl |= kFlagNee;
            else st.srays[4 * size_t(i)].w = 0.f;
        }
        st.rays[2 * size_t(i)].w = 0.f;
        st.flags[i] = (fl & ~kDepthMask) | ((depth + 1) & kDepthMask);
        st.L[i] = s4f(L);
        st.rng[i] = smp.rng.state;
        st.dim[i] = smp.dim;
        return;
    }
    const f3 wo_n = norm3(neg3(rd));
    const uint32_t light = info.y & 0x7FFFFFFFu;
    if (light) add_emission(prim, light, n, rd);
    fl = (fl & ~kDepthMask) | ((depth + 1) & kDepthMask);
    // 3. material: BSDF (interaction.cu:77-89, bsdf.cu), sample_ld, sample_f
    const DMat m = S.mats[info.x];
    const f3 fx = norm3(dpdu);
    const f3 fy = cross3(ns, fx);
    Bxdf bx;
    bx.type = m.type;
    bx.tr = TR{m.ax, m.ay};
    if (!bx.tr.smooth()) bx.tr = TR{smax(m.ax, 1e-4f), smax(m.ay, 1e-4f)};
    bx.r = s4(0.f);
    bx.eta = s4(0.f);
    bx.k = s4(0.f);
    bx.eta_d = 1.f;
    if (m.type == kDiffuse) {
        for (int k = 0; k < kNS; ++k) bx.r.v[k] = clampf(sigmoid_eval(m.c0, m.c1, m.c2, lam.v[k]), 0.0f, 1.0f);
        if (GRAD) {
            // every later contribution is linear in this vertex's reflectance R:
            // dLoss/dR_k = g_k (Ltot_k - L_k) / R_k, R = sigmoid(c0 l^2 + c1 l + c2)
            const S4 lt = f4s(P.gd.ltot[i]);
            float d0 = 0.f, d1 = 0.f, d2 = 0.f;
#pragma unroll
            for (int k = 0; k < kNS; ++k) {
                const float r = bx.r.v[k];
                if (!(r > 0.f && r < 1.f)) continue;
                const float l = lam.v[k];
                const float x = fmaf(l, fmaf(l, m.c0, m.c1), m.c2);
                const float q = 1.f + x * x;
                const float v = gadj.v[k] * (lt.v[k] - L.v[k]) / r * (0.5f / (q * sqrtf(q)));
                d0 += v * l * l;
                d1 += v * l;
                d2 += v;
            }
            *grec = make_float4(d0, d1, d2, __uint_as_float(info.x));  // added by the kernel, wave-aggregated
        }
    } else if (m.type == kConductor) {
        for (int k = 0; k < kNS; ++k) {
            bx.eta.v[k] = spec_eval(S, m.eta_off, m.eta_n, lam.v[k]);
            bx.k.v[k] = spec_eval(S, m.k_off, m.k_n, lam.v[k]);
        }
    } else {  // materials/dielectric_material.cu:37-70
        float e = m.eta_c;
        if (m.eta_n) {
            e = spec_eval(S, m.eta_off, m.eta_n, lam.v[0]);
            if (!(lpdf.v[1] == 0 && lpdf.v[2] == 0 && lpdf.v[3] == 0)) {  // terminate_secondary
                lpdf.v[1] = lpdf.v[2] = lpdf.v[3] = 0;
                lpdf.v[0] /= kNS;
                st.lam_pdf[i] = s4f(lpdf);
            }
        }
        bx.eta_d = e == 0 ? 1.f : e;
    }
    const uint32_t bflags = bx.flags();
    bool nee = false;
    if (bflags & (4u | 8u)) {
        // sample_ld (wavefront_path.cu:574-624)
        P3i cpi = ipi;
        const bool refl = (bflags & 1u) != 0, trans = (bflags & 2u) != 0;
        if (refl && !trans) cpi = p3i_exact(offset_ray_origin(ipi, n, wo_n));
        else if (trans && !refl) cpi = p3i_exact(offset_ray_origin(ipi, n, neg3(wo_n)));
        const float u = smp.get_1d();
        float ul0, ul1;
        smp.get_2d(&ul0, &ul1);
        if (S.num_lights > 0) {
            const uint32_t li = search_cdf(u, S.light_cdf, S.num_lights);
            const float pmf = S.light_pmf[li];
            const DLight Lg = S.lights[li];
            const f3 cpt = p3i_mid(cpi);
            ShapeSample ss;
            f3 wi = {0.f, 0.f, 0.f};
            S4 Le = s4(0.f);
            bool ok;
            if (Lg.tri == kInfiniteLight) {
                // UniformInfiniteLight::sample_li (uniform_infinite_light.cu:51-60): the
                // light point is ctx.p() + wi 2R, exact and without a normal
                wi = uniform_sphere(ul0, ul1);
                ss.pdf = kUniformSpherePdf;
                ss.pi = p3i_exact(add3(cpt, mul3(S.inf_dist, wi)));
                ss.n = f3{0.f, 0.f, 0.f};
                Le = light_L(S, Lg, ss.n, ss.n, lam);
                ok = s4pos(Le);
            } else {
                ok = prim_light_sample(S, Lg.tri, cpi, n, ns, ul0, ul1, &ss);
                const f3 sp = p3i_mid(ss.pi);
                ok = ok && !(ss.pdf == 0 || sqlen3(sub3(sp, cpt)) == 0);
                if (ok) {
                    wi = norm3(sub3(sp, cpt));
                    Le = light_L(S, Lg, ss.n, neg3(wi), lam);
                    ok = s4pos(Le) && ss.pdf != 0;
                }
            }
            {
                if (ok) {
                    // BSDF::f and BSDF::pdf (bsdf.cu:44-93) in the shading frame
                    const f3 wil = {dot3(wi, fx), dot3(wi, fy), dot3(wi, ns)};
                    const f3 wol = {dot3(wo_n, fx), dot3(wo_n, fy), dot3(wo_n, ns)};
                    S4 f = wol.z == 0 ? s4(0.f) : bx.f(wol, wil);
                    f = s4scale(f, absdot3(wi, ns));
                    if (s4pos(f)) {
                        const float pdf_b = wol.z == 0 ? 0.f : bx.pdf(wol, wil);
                        const float pdf_light = pmf * ss.pdf;
                        const float wl = power_heuristic(pdf_light, pdf_b);
                        const S4 Ld = s4div(s4mul(s4scale(Le, wl), f), pdf_light);
                        st.nee_L[i] = s4f(s4mul(beta, Ld));
                        nee = true;
                        if (GRAD)
                            P.gd.nee[i] = make_float4(dot4(gadj, s4mul(beta, Ld)), wl, __uint_as_float(Lg.group), 0.f);
                        f3 o, d;
                        spawn_ray_to(ipi, n, ss.pi, ss.n, &o, &d);
                        store_ray(st.srays + 4 * size_t(i), o, 0.6f, d);
                        spawn_ray_to(ss.pi, ss.n, ipi, n, &o, &d);
                        store_ray(st.srays + 4 * size_t(i) + 2, o, 0.6f, d);
                    }
                }
            }
        }
    }
    if (nee) fl |= kFlagNee;
    else {
        st.srays[4 * size_t(i)].w = 0.f;
        st.srays[4 * size_t(i) + 2].w = 0.f;
    }
    // BSDF::sample_f (bsdf.cu:57-80)
    const f3 wo = neg3(rd);
    const float uc = smp.get_1d();
    float u0, u1;
    smp.get_2d(&u0, &u1);
    const f3 wol = {dot3(wo, fx), dot3(wo, fy), dot3(wo, ns)};
    BxSample bs;
    if (!bxdf_sample(bx, wol, uc, u0, u1, &bs)) {
        beta = s4(0.0f);  // the stale ray is traced again and the path ends next round
    } else {
        const f3 wi = add3(add3(mul3(bs.wi.x, fx), mul3(bs.wi.y, fy)), mul3(bs.wi.z, ns));
        beta = s4mul(beta, s4div(s4scale(bs.f, absdot3(wi, ns)), bs.pdf));
        scal.y = bs.pdf;
        fl = bs.specular ? (fl | kFlagSpecular) : (fl & ~kFlagSpecular);
        if (!bs.specular) fl |= kFlagAnyNonSpec;
        if (bs.transmission) scal.z *= sqrf(bs.eta);
        st.prev[3 * size_t(i)] = make_float4(ipi.x.lo, ipi.x.hi, ipi.y.lo, ipi.y.hi);
        st.prev[3 * size_t(i) + 1] = make_float4(ipi.z.lo, ipi.z.hi, n.x, n.y);
        st.prev[3 * size_t(i) + 2] = make_float4(n.z, ns.x, ns.y, ns.z);
        store_ray(st.rays + 2 * size_t(i), offset_ray_origin(ipi, n, wi), __builtin_inff(), wi);
    }
    st.flags[i] = fl;
    st.L[i] = s4f(L);
    st.beta[i] = s4f(beta);
    st.scal[i] = scal;
    st.rng[i] = smp.rng.state;
    st.dim[i] = smp.dim;
}

// The path-integrator shading instances need about 172 VGPRs unbounded (2 waves
// per SIMD); held to 168 (3 waves) they spill 8 bytes per lane, which is faster.
#define PBRTM_SHADE_WAVES __attribute__((amdgpu_waves_per_eu(3)))

template <uint32_t MODE>
__global__ __launch_bounds__(kBlock) PBRTM_SHADE_WAVES void shade_kernel(SceneDev S, State st, Params P) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pid = i < P.pool ? st.path_id[i] : kIdle;
    shade_impl<false, false, MODE>(S, st, P, i, pid, pid, P.spd, P.max_depth, P.round, nullptr, nullptr);
}

// The gradient replay of one round (same paths as the forward render).
constexpr uint32_t kMaxGroups = 255;
__global__ __launch_bounds__(kBlock) void shade_grad_kernel(SceneDev S, State st, Params P) {
    __shared__ double s_acc[kMaxGroups + 1];
    for (uint32_t k = threadIdx.x; k <= P.gd.ngroups; k += kBlock) s_acc[k] = 0.0;
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pid = i < P.pool ? st.path_id[i] : kIdle;
    float4 grec = make_float4(0.f, 0.f, 0.f, __uint_as_float(kIdle));
    float4 erec[2] = {make_float4(0.f, 0.f, __uint_as_float(kIdle), 0.f),
                      make_float4(0.f, 0.f, __uint_as_float(kIdle), 0.f)};
    shade_impl<false, true>(S, st, P, i, pid, pid, P.spd, P.max_depth, P.round, erec, &grec);
    // Emitter sums: the pmf part over the whole wave, the scale part per emitter
    // (lanes of one emitter summed first), into shared memory.
    {
        double kk = double(erec[0].y) + double(erec[1].y);
        for (int d = kWave / 2; d > 0; d >>= 1) kk += __shfl_xor(kk, d, kWave);
        if (lane_id() == 0 && kk != 0.0) atomicAdd(&s_acc[P.gd.ngroups], kk);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const uint32_t grp = __float_as_uint(erec[e].z);
            bool act = grp != kIdle && erec[e].x != 0.f;
            for (int r = 0; r < 4; ++r) {
                const uint64_t pend = __ballot(act);
                if (!pend) break;
                const int leader = __ffsll((unsigned long long)pend) - 1;
                const uint32_t g0 = __shfl(grp, leader, kWave);
                const bool mine = act && grp == g0;
                double v = mine ? double(erec[e].x) : 0.0;
                for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
                if (lane_id() == leader) atomicAdd(&s_acc[g0], v);
                if (mine) act = false;
            }
            if (act) atomicAdd(&s_acc[grp], double(erec[e].x));
        }
    }
    // Coefficient derivatives: lanes of a wave that hit the same material (walls,
    // large triangles) are summed first, for up to three materials per wave, the
    // rest add directly; each block adds into one of kGradShards copies of the
    // table (summed afterwards), so that the atomics on a few hot materials from
    // every wave of the launch do not serialise on one address.
    float *shard = P.gd.d_coef + size_t(blockIdx.x % kGradShards) * P.gd.ncoef;
    uint32_t mat = __float_as_uint(grec.w);
    bool act = mat != kIdle && (grec.x != 0.f || grec.y != 0.f || grec.z != 0.f);
    for (int r = 0; r < 3; ++r) {
        const uint64_t pend = __ballot(act);
        if (!pend) break;
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const uint32_t m0 = __shfl(mat, leader, kWave);
        const bool mine = act && mat == m0;
        float a = mine ? grec.x : 0.f, b = mine ? grec.y : 0.f, c = mine ? grec.z : 0.f;
        for (int d = kWave / 2; d > 0; d >>= 1) {
            a += __shfl_xor(a, d, kWave);
            b += __shfl_xor(b, d, kWave);
            c += __shfl_xor(c, d, kWave);
        }
        if (lane_id() == leader) {
            float *dc = shard + 3 * size_t(m0);
            atomicAdd(dc, a);
            atomicAdd(dc + 1, b);
            atomicAdd(dc + 2, c);
        }
        if (mine) act = false;
    }
    if (act) {
        float *dc = shard + 3 * size_t(mat);
        atomicAdd(dc, grec.x);
        atomicAdd(dc + 1, grec.y);
        atomicAdd(dc + 2, grec.z);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k <= P.gd.ngroups; k += kBlock)
        if (s_acc[k] != 0.0) atomicAdd(&P.gd.acc[k], s_acc[k]);
}

// Batched render: each slot reads its job's scene and parameters; the round
// number comes from the device counter so the round can be replayed in a graph.
__global__ __launch_bounds__(kBlock) PBRTM_SHADE_WAVES void shade_multi_kernel(State st, Params P) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pid = i < P.pool ? st.path_id[i] : kIdle;
    const uint32_t j = pid != kIdle ? st.job[i] : 0u;
    const JobDev &J = P.jobs[j];
    const uint32_t lid = pid != kIdle ? uint32_t(pid - J.base) : kIdle;
    shade_impl<true, false>(J.S, st, P, i, pid, lid, J.spd, J.max_depth, st.counters[kCntRound], nullptr,
                            nullptr);
}

// Sums the kGradShards copies of the coefficient-derivative table into the first.
__global__ __launch_bounds__(kBlock) void reduce_shards_kernel(float *d, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float v = d[i];
    for (uint32_t k = 1; k < kGradShards; ++k) v += d[size_t(k) * n + i];
    d[i] = v;
}

// Start and end of a batched round (kernels, so that a captured graph holds no memset).
__global__ void round_begin_kernel(uint32_t *counters) {
    if (threadIdx.x == 0) {
        counters[kCntQueue] = 0;
        counters[kCntSQueue] = 0;
    }
}
// A round that traced no ray had no live path: the render is over and the round
// does not count.
__global__ void round_end_kernel(uint32_t *counters) {
    if (threadIdx.x != 0 || counters[kCntDone]) return;
    if (counters[kCntQueue] == 0) {
        counters[kCntDone] = 1;
        return;
    }
    unsigned long long *acc = reinterpret_cast<unsigned long long *>(counters + kCntAccRays);
    acc[0] += counters[kCntQueue];
    acc[1] += counters[kCntSQueue];
    counters[kCntRound] += 1;
}

// The reference's one-by-one pool schedule for one job (one workgroup per
// job), recovered from each path's length in rounds: ids [0, pool) start in
// round 0; the ids started at the end of round r are the next T_r ids, T_r =
// the number of paths finishing in round r (wavefront_path.cu:265-285,
// :641-700). rec_round holds lengths on entry and finishing rounds on exit;
// rounds[j] receives the job's round count. Lengths never exceed
// max_depth + 1, so finishing counts live in a ring of kRing rounds.
constexpr uint32_t kSchedBlock = 1024, kRing = 4096;
constexpr int kSchedPaths = 8;  // paths per thread and trip
__global__ __launch_bounds__(kSchedBlock) void schedule_kernel(const JobDev *jobs, uint32_t *rec_round,
                                                               uint32_t *rounds, uint32_t *err) {
    __shared__ uint32_t ring[kRing];
    __shared__ uint32_t s_max, s_bad;
    const JobDev &J = jobs[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < kRing; k += kSchedBlock) ring[k] = 0;
    if (threadIdx.x == 0) {
        s_max = 0;
        s_bad = 0;
    }
    __syncthreads();
    uint32_t *rr = rec_round + J.base;
    const uint64_t total = J.total;
    const uint32_t lmax = J.max_depth + 1;
    const uint64_t cap = (total + 1) * (uint64_t(J.max_depth) + 3);
    uint64_t lo = 0, hi = total < J.vpool ? total : J.vpool;
    uint64_t r = 0;
    uint32_t my_max = 0;
    for (;;) {
        // wave-uniform trips of kSchedPaths paths per thread (their lengths loaded
        // together); the finishing rounds of a wave's paths are counted per distinct
        // value (at most max_depth + 1 of them), one LDS atomic each
        for (uint64_t b = lo; b < hi; b += uint64_t(kSchedBlock) * kSchedPaths) {
            uint32_t len[kSchedPaths];
#pragma unroll
            for (int j = 0; j < kSchedPaths; ++j) {
                const uint64_t k = b + uint64_t(j) * kSchedBlock + threadIdx.x;
                len[j] = k < hi ? rr[k] : 0u;
            }
#pragma unroll
            for (int j = 0; j < kSchedPaths; ++j) {
                const uint64_t k = b + uint64_t(j) * kSchedBlock + threadIdx.x;
                bool ok = k < hi;
                if (ok && len[j] > lmax) {
                    s_bad = 1;
                    ok = false;
                }
                const uint32_t f = uint32_t(r) + len[j];
                if (ok) {
                    rr[k] = f;
                    my_max = f > my_max ? f : my_max;
                }
                uint64_t pending = __ballot(ok);
                while (pending) {
                    const int leader = __ffsll((unsigned long long)pending) - 1;
                    const uint32_t fl = __shfl(f, leader, kWave);
                    const uint64_t same = __ballot(ok && f == fl);
                    if (lane_id() == uint32_t(leader)) atomicAdd(&ring[fl % kRing], uint32_t(__popcll(same)));
                    pending &= ~same;
                }
            }
        }
        __syncthreads();
        const uint32_t t = ring[r % kRing];
        __syncthreads();
        if (threadIdx.x == 0) ring[r % kRing] = 0;
        if (hi == total || s_bad) break;
        lo = hi;
        hi = hi + t < total ? hi + t : total;
        if (++r > cap) {
            if (threadIdx.x == 0) s_bad = 1;
            break;
        }
    }
    atomicMax(&s_max, my_max);
    __syncthreads();
    if (threadIdx.x == 0) {
        rounds[blockIdx.x] = s_max + 1;
        if (s_bad) atomicOr(err, uint32_t(kErrCapacity));
    }
}

// write_frame_buffer + RGBFilm::add_sample/get_pixel_rgb (rgb_film.cu:85-144):
// one wave per pixel ranks its samples by (round, sample), sums them in that
// order, divides by the weight sum and converts sensor RGB to output RGB with
// the compensated 3x3 product.
__device__ __forceinline__ float ip3(const float *m, const float *v) {
    float vv = m[2] * v[2];
    float err = fmaf(m[2], v[2], -vv);
    for (int k = 1; k >= 0; --k) {
        const float ab = m[k] * v[k];
        const float ab_err = fmaf(m[k], v[k], -ab);
        const float s = ab + vv;
        const float delta = s - ab;
        const float s_err = (ab - (s - delta)) + (vv - delta);
        err = ab_err + (err + s_err);
        vv = s;
    }
    return vv + err;
}

constexpr uint32_t kFilmWaves = 4;
constexpr uint32_t kMaxSpp = 1024;

__global__ __launch_bounds__(kFilmWaves * kWave) void film_kernel(SceneDev S, const float4 *__restrict__ rec,
                                                                  const uint32_t *__restrict__ rec_round, uint32_t spp,
                                                                  float *film_rgb, float *film_aux) {
    __shared__ uint32_t s_key[kFilmWaves][kMaxSpp];
    __shared__ uint16_t s_order[kFilmWaves][kMaxSpp];
    __shared__ float4 s_rec[kFilmWaves][kWave];
    const uint32_t wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint32_t wh = S.width * S.height;
    const uint32_t px = blockIdx.x * kFilmWaves + wv;
    if (px >= wh) return;
    for (uint32_t s = lane; s < spp; s += kWave) s_key[wv][s] = rec_round[size_t(s) * wh + px];
    __builtin_amdgcn_wave_barrier();
    for (uint32_t s = lane; s < spp; s += kWave) {
        const uint32_t k = s_key[wv][s];
        uint32_t rank = 0;
        for (uint32_t t = 0; t < spp; ++t) {
            const uint32_t kt = s_key[wv][t];
            rank += (kt < k || (kt == k && t < s)) ? 1u : 0u;
        }
        s_order[wv][rank] = uint16_t(s);
    }
    __builtin_amdgcn_wave_barrier();
    // The records are gathered 64 at a time by all lanes (each from its own sample
    // row), then lane 0 adds them in rank order from LDS.
    float sum[3] = {0.f, 0.f, 0.f}, wsum = 0.f;
    for (uint32_t c = 0; c < spp; c += kWave) {
        if (c + lane < spp) s_rec[wv][lane] = rec[size_t(s_order[wv][c + lane]) * wh + px];
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            const uint32_t m = spp - c < kWave ? spp - c : kWave;
            for (uint32_t j = 0; j < m; ++j) {
                const float4 r = s_rec[wv][j];
                if (r.w == 0) continue;
                sum[0] += r.w * r.x;
                sum[1] += r.w * r.y;
                sum[2] += r.w * r.z;
                wsum += r.w;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane != 0) return;
    float rgb[3] = {sum[0], sum[1], sum[2]};
    if (wsum != 0)
        for (int c = 0; c < 3; ++c) rgb[c] /= wsum;
    for (int c = 0; c < 3; ++c) film_rgb[3 * size_t(px) + c] = ip3(S.out_from_sensor + 3 * c, rgb);
    if (film_aux) {
        for (int c = 0; c < 3; ++c) film_aux[4 * size_t(px) + c] = sum[c];
        film_aux[4 * size_t(px) + 3] = wsum;
    }
}

// film_kernel for at most kLaneFilm samples per pixel: one lane per pixel (a wave
// per pixel would leave most lanes idle), the same (round, sample) order and the
// same sums, one operation at a time.
constexpr uint32_t kLaneFilm = 4;

__global__ __launch_bounds__(kBlock) void film_lane_kernel(SceneDev S, const float4 *__restrict__ rec,
                                                           const uint32_t *__restrict__ rec_round, uint32_t spp,
                                                           float *film_rgb, float *film_aux) {
    const uint32_t wh = S.width * S.height;
    const uint32_t px = blockIdx.x * kBlock + threadIdx.x;
    if (px >= wh) return;
    uint32_t key[kLaneFilm];
    float4 r[kLaneFilm];
#pragma unroll
    for (uint32_t s = 0; s < kLaneFilm; ++s) {
        key[s] = 0u;
        r[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (s < spp) {
            key[s] = rec_round[size_t(s) * wh + px];
            r[s] = rec[size_t(s) * wh + px];
        }
    }
    float sum[3] = {0.f, 0.f, 0.f}, wsum = 0.f;
#pragma unroll
    for (uint32_t pos = 0; pos < kLaneFilm; ++pos) {
        if (pos >= spp) break;
        float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (uint32_t s = 0; s < kLaneFilm; ++s) {
            uint32_t rank = 0;
#pragma unroll
            for (uint32_t t = 0; t < kLaneFilm; ++t)
                rank += (t < spp && (key[t] < key[s] || (key[t] == key[s] && t < s))) ? 1u : 0u;
            if (s < spp && rank == pos) q = r[s];
        }
        if (q.w == 0) continue;
        sum[0] += q.w * q.x;
        sum[1] += q.w * q.y;
        sum[2] += q.w * q.z;
        wsum += q.w;
    }
    float rgb[3] = {sum[0], sum[1], sum[2]};
    if (wsum != 0)
        for (int c = 0; c < 3; ++c) rgb[c] /= wsum;
    for (int c = 0; c < 3; ++c) film_rgb[3 * size_t(px) + c] = ip3(S.out_from_sensor + 3 * c, rgb);
    if (film_aux) {
        for (int c = 0; c < 3; ++c) film_aux[4 * size_t(px) + c] = sum[c];
        film_aux[4 * size_t(px) + 3] = wsum;
    }
}

// The film of one render: a lane per pixel up to kLaneFilm samples, else a wave.
static void launch_film(const SceneDev &S, const float4 *rec, const uint32_t *rec_round, uint32_t spp,
                        float *film_rgb, float *film_aux, hipStream_t stream) {
    const uint32_t wh = S.width * S.height;
    if (spp <= kLaneFilm)
        film_lane_kernel<<<(wh + kBlock - 1) / kBlock, kBlock, 0, stream>>>(S, rec, rec_round, spp, film_rgb,
                                                                            film_aux);
    else
        film_kernel<<<(wh + kFilmWaves - 1) / kFilmWaves, kFilmWaves * kWave, 0, stream>>>(S, rec, rec_round, spp,
                                                                                         film_rgb, film_aux);
}

}  // namespace

namespace {

// Per pixel: loss terms (film - target)^2 and the adjoint of rgb_sum / weight_sum.
__global__ __launch_bounds__(kBlock) void film_adjoint_kernel(SceneDev S, const float *__restrict__ film,
                                                              const float *__restrict__ aux,
                                                              const float *__restrict__ target,
                                                              float *__restrict__ adj, double *loss) {
    __shared__ double s_sum[kBlock / kWave];
    const uint32_t px = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t wh = S.width * S.height;
    double l = 0.0;
    if (px < wh) {
        float e[3];
        for (int k = 0; k < 3; ++k) {
            e[k] = film[3 * size_t(px) + k] - target[3 * size_t(px) + k];
            l += double(e[k]) * double(e[k]);
        }
        const float w = aux[4 * size_t(px) + 3];
        for (int j = 0; j < 3; ++j) {
            float v = 0.f;
            for (int k = 0; k < 3; ++k) v += 2.f * e[k] * S.out_from_sensor[3 * k + j];
            adj[3 * size_t(px) + j] = w != 0.f ? v / w : v;
        }
    }
    for (int d = kWave / 2; d > 0; d >>= 1) l += __shfl_xor(l, d, kWave);
    if (lane_id() == 0) s_sum[threadIdx.x / kWave] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t k = 0; k < kBlock / kWave; ++k) t += s_sum[k];
        atomicAdd(loss, t);
    }
}

}  // namespace

hipError_t film_adjoint(const SceneDev &S, const float *film_rgb, const float *film_aux, const float *target,
                        float *adj, double *loss, hipStream_t stream) {
    const uint32_t wh = S.width * S.height;
    film_adjoint_kernel<<<(wh + kBlock - 1) / kBlock, kBlock, 0, stream>>>(S, film_rgb, film_aux, target, adj, loss);
    return hipGetLastError();
}

namespace {

hipError_t reserve_pool(PathPool &pp, uint32_t pool, uint64_t total, bool batch) {
    hipError_t e;
#define PT_CK(x)                          \
    do {                                  \
        if ((e = (x)) != hipSuccess) return e; \
    } while (0)
    if (pp.cap < pool) {
        PT_CK(pp.path_id.reserve(size_t(pool) * 4));
        PT_CK(pp.flags.reserve(size_t(pool) * 4));
        PT_CK(pp.dim.reserve(size_t(pool) * 4));
        PT_CK(pp.rng.reserve(size_t(pool) * 8));
        for (DevBuf *b : {&pp.lam, &pp.lam_pdf, &pp.L, &pp.beta, &pp.nee_L, &pp.scal}) PT_CK(b->reserve(size_t(pool) * 16));
        PT_CK(pp.prev.reserve(size_t(pool) * 48));
        PT_CK(pp.rays.reserve(size_t(pool) * 32));
        PT_CK(pp.prim.reserve(size_t(pool) * 4));
        PT_CK(pp.thit.reserve(size_t(pool) * 4));
        PT_CK(pp.bary.reserve(size_t(pool) * 12));
        PT_CK(pp.srays.reserve(size_t(pool) * 64));
        PT_CK(pp.sprim.reserve(size_t(pool) * 8));
        PT_CK(pp.queue.reserve(size_t(pool) * 4));
        PT_CK(pp.squeue.reserve(size_t(pool) * 8));
        pp.cap = pool;
    }
    if (batch) {
        PT_CK(pp.job.reserve(size_t(pool) * 4));
        PT_CK(pp.sround.reserve(size_t(pool) * 4));
    }
    if (pp.rec_cap < total) {
        PT_CK(pp.rec.reserve(total * 16));
        PT_CK(pp.rec_round.reserve(total * 4));
        pp.rec_cap = total;
    }
    PT_CK(pp.counters.reserve(64));
    return hipSuccess;
#undef PT_CK
}

State make_state(PathPool &pp, uint32_t *err) {
    State st{};
    st.path_id = pp.path_id.as<uint32_t>();
    st.flags = pp.flags.as<uint32_t>();
    st.dim = pp.dim.as<uint32_t>();
    st.rng = pp.rng.as<uint64_t>();
    st.lam = pp.lam.as<float4>();
    st.lam_pdf = pp.lam_pdf.as<float4>();
    st.L = pp.L.as<float4>();
    st.beta = pp.beta.as<float4>();
    st.nee_L = pp.nee_L.as<float4>();
    st.scal = pp.scal.as<float4>();
    st.prev = pp.prev.as<float4>();
    st.rays = pp.rays.as<float4>();
    st.prim = pp.prim.as<uint32_t>();
    st.thit = pp.thit.as<float>();
    st.bary = pp.bary.as<float>();
    st.srays = pp.srays.as<float4>();
    st.sprim = pp.sprim.as<uint32_t>();
    st.rec = pp.rec.as<float4>();
    st.rec_round = pp.rec_round.as<uint32_t>();
    st.job = pp.job.as<uint32_t>();
    st.sround = pp.sround.as<uint32_t>();
    st.lfin = nullptr;
    st.pfin = nullptr;
    st.counters = pp.counters.as<uint32_t>();
    st.err = err;
    return st;
}

}  // namespace

__global__ __launch_bounds__(256) void tag_leaf_kernel(float *__restrict__ tri_leaf,
                                                       const uint32_t *__restrict__ prim_order,
                                                       const uint32_t *__restrict__ prim_shape, uint32_t n) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = prim_order[k];
    tri_leaf[12 * size_t(k) + 9] = __uint_as_float(p < n ? prim_shape[p] : 0u);
}

hipError_t tag_leaf_shapes(float *tri_leaf, const uint32_t *prim_order, const uint32_t *prim_shape, uint32_t n,
                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    tag_leaf_kernel<<<(n + 255) / 256, 256, 0, stream>>>(tri_leaf, prim_order, prim_shape, n);
    return hipGetLastError();
}

hipError_t render(const SceneDev &S, const WideTrace &W0, uint32_t spp, uint32_t max_depth, uint32_t pool,
                  PathPool &pp, float *film_rgb, float *film_aux, uint32_t *err, RenderStats *stats,
                  hipStream_t stream, float4 *lfin, float4 *pfin, const GradDev *gd, uint32_t integrator,
                  unsigned long long *counts) {
    if (integrator > kIntSurfaceNormal || (gd && integrator != kIntPath)) return hipErrorInvalidValue;
    const uint64_t total = uint64_t(S.width) * S.height * spp;
    const uint32_t spd = uint32_t(sqrt(double(spp)));
    if (spd * spd != spp || spp == 0 || spp > kMaxSpp) return hipErrorInvalidValue;
    // the new-path counter runs up to total + pool
    if (total + pool >= 0xFFFFFFFFull || pool == 0) return hipErrorInvalidValue;
    hipError_t e;
#define PT_CK(x)                          \
    do {                                  \
        if ((e = (x)) != hipSuccess) return e; \
    } while (0)
    PT_CK(reserve_pool(pp, pool, total, false));
    PT_CK(reserve_ray_sort(pp.rsort, 2 * pool));  // sized once: no reallocation while rounds are in flight
    WideTrace W = W0;
    const uint32_t spill_need = W.spill_per_ray;
    if (spill_need) {
        PT_CK(pp.spill.reserve(size_t(pool) * 2 * spill_need * sizeof(uint2)));
        W.spill = pp.spill.as<uint2>();
    }
    State st = make_state(pp, err);
    st.lfin = lfin;
    st.pfin = pfin;
    Params P{total, spd, spp, max_depth, 0u, pool, nullptr, nullptr, 0u};
    if (gd) {
        if (gd->ngroups > kMaxGroups) return hipErrorInvalidValue;
        P.gd = *gd;
    }
    const uint32_t grid = (pool + kBlock - 1) / kBlock;
    // ev[0]: start, ev[1]: end of the rounds, ev[2]: end; per round slot (rounds
    // alternate between two): before the closest-hit trace, after it, after the
    // shadow trace, after shading, after the counter read-back
    constexpr int kEv = 3 + 2 * 5;
    hipEvent_t ev[kEv];
    for (auto &x : ev) PT_CK(hipEventCreate(&x));
    uint32_t *hc = nullptr;  // pinned: the counters of two rounds
    PT_CK(hipHostMalloc(reinterpret_cast<void **>(&hc), 2 * 16 * sizeof(uint32_t), hipHostMallocDefault));
    struct EvGuard {
        hipEvent_t *e;
        uint32_t *h;
        hipStream_t s;
        ~EvGuard() {
            hipStreamSynchronize(s);  // no read-back may still target h
            for (int k = 0; k < kEv; ++k) hipEventDestroy(e[k]);
            hipHostFree(h);
        }
    } guard{ev, hc, stream};
    PT_CK(hipMemsetAsync(st.counters, 0, 64, stream));
    uint32_t first = uint32_t(std::min<uint64_t>(total, pool));
    PT_CK(hipMemcpyAsync(st.counters + kCntNextId, &first, 4, hipMemcpyHostToDevice, stream));
    uint64_t next_before = first;  // paths started before the current round
    PT_CK(hipEventRecord(ev[0], stream));
    if (gd) init_pool_kernel<true><<<grid, kBlock, 0, stream>>>(S, st, P);
    else init_pool_kernel<false><<<grid, kBlock, 0, stream>>>(S, st, P);
    PT_CK(hipGetLastError());
    RenderStats rs;
    float t_trace = 0, t_shadow = 0, t_shade = 0;
    struct {
        bool sort_closest = true, sort_shadow = true;
        double ms[2][2] = {{0, 0}, {0, 0}};
        uint64_t rays[2][2] = {{0, 0}, {0, 0}};
        uint32_t quarters = 4;
        double occ_ms[3] = {0, 0, 0};
        uint64_t occ_rays[3] = {0, 0, 0};
    } tune;
    // One round: both traces and the shading kernel, then the counters to the host.
    // Round r + 1 is enqueued before round r's counters are read, so the device
    // does not wait for the host between rounds; when round r turns out to be the
    // last, round r + 1 finds no live path and does nothing.
    auto enqueue = [&](uint32_t round) -> hipError_t {
        hipError_t e;
        hipEvent_t *E = ev + 3 + 5 * (round & 1);
        P.round = round;
        PT_CK(hipMemsetAsync(st.counters + kCntQueue, 0, 8, stream));
        PT_CK(hipEventRecord(E[0], stream));
        const bool tuning = round >= kTuneFirst && round < kTuneFirst + 4;
        const bool sort_c = tuning ? ((round - kTuneFirst) % 2 == 0) : tune.sort_closest;
        const bool sort_s = tuning ? ((round - kTuneFirst) % 2 == 0) : tune.sort_shadow;
        const bool occ_tuning = round >= kOccFirst && round < kOccFirst + 6;
        W.grid_quarters = occ_tuning ? kOccQuarters[(round - kOccFirst) % 3] : tune.quarters;
        if (sort_c) {  // closest-hit rays in (octant, origin Morton code) order
            uint32_t *q = nullptr;
            PT_CK(sort_rays(W, reinterpret_cast<const float *>(st.rays), pool, pp.prim.as<uint32_t>(),
                            pp.thit.as<float>(), pp.bary.as<float>(), err, pp.rsort, st.counters + kCntQueue, &q,
                            stream));
            PT_CK(trace_wide_queued(W, reinterpret_cast<const float *>(st.rays), pool, 0, pp.prim.as<uint32_t>(),
                                    pp.thit.as<float>(), pp.bary.as<float>(), err, q, st.counters + kCntQueue,
                                    stream, counts));
        } else {
            PT_CK(trace_wide(W, reinterpret_cast<const float *>(st.rays), pool, 0, pp.prim.as<uint32_t>(),
                             pp.thit.as<float>(), pp.bary.as<float>(), err, counts, pp.queue.as<uint32_t>(),
                             st.counters + kCntQueue, stream));
        }
        PT_CK(hipEventRecord(E[1], stream));
        if (spill_need) W.spill = pp.spill.as<uint2>();
        if (sort_s) {
            uint32_t *q = nullptr;
            PT_CK(sort_rays(W, reinterpret_cast<const float *>(st.srays), 2 * pool, pp.sprim.as<uint32_t>(), nullptr,
                            nullptr, err, pp.rsort, st.counters + kCntSQueue, &q, stream));
            PT_CK(trace_wide_queued(W, reinterpret_cast<const float *>(st.srays), 2 * pool, 1,
                                    pp.sprim.as<uint32_t>(), nullptr, nullptr, err, q, st.counters + kCntSQueue,
                                    stream, counts ? counts + 3 : nullptr));
        } else {
            PT_CK(trace_wide(W, reinterpret_cast<const float *>(st.srays), 2 * pool, 1, pp.sprim.as<uint32_t>(),
                             nullptr, nullptr, err, counts ? counts + 3 : nullptr, pp.squeue.as<uint32_t>(),
                             st.counters + kCntSQueue, stream));
        }
        PT_CK(hipEventRecord(E[2], stream));
        if (gd) shade_grad_kernel<<<grid, kBlock, 0, stream>>>(S, st, P);
        else if (integrator == kIntMegakernelPath) shade_kernel<kIntMegakernelPath><<<grid, kBlock, 0, stream>>>(S, st, P);
        else if (integrator == kIntAmbientOcclusion)
            shade_kernel<kIntAmbientOcclusion><<<grid, kBlock, 0, stream>>>(S, st, P);
        else if (integrator == kIntSurfaceNormal) shade_kernel<kIntSurfaceNormal><<<grid, kBlock, 0, stream>>>(S, st, P);
        else shade_kernel<kIntPath><<<grid, kBlock, 0, stream>>>(S, st, P);
        PT_CK(hipGetLastError());
        PT_CK(hipEventRecord(E[3], stream));
        PT_CK(hipMemcpyAsync(hc + 16 * (round & 1), st.counters, 32, hipMemcpyDeviceToHost, stream));
        PT_CK(hipEventRecord(E[4], stream));
        return hipSuccess;
    };
    PT_CK(enqueue(0));
    uint32_t r = 0;
    for (;;) {
        PT_CK(enqueue(r + 1));
        hipEvent_t *E = ev + 3 + 5 * (r & 1);
        PT_CK(hipEventSynchronize(E[4]));
        const uint32_t *host_cnt = hc + 16 * (r & 1);
        // Done when no ray was traced and no path could start: every live slot then
        // ended this round. A wavefront path always has a ray, so there this round
        // had no live path and does not count; the megakernel integrators' single-
        // vertex paths retire their ray a round before they end, and that last
        // round's shadow rays and film records do count.
        const bool done = host_cnt[kCntQueue] == 0 && next_before >= total;
        if (done && integrator == kIntPath) break;
        next_before = host_cnt[kCntNextId];
        float a, b, c;
        hipEventElapsedTime(&a, E[0], E[1]);
        hipEventElapsedTime(&b, E[1], E[2]);
        hipEventElapsedTime(&c, E[2], E[3]);
        t_trace += a;
        t_shadow += b;
        t_shade += c;
        if (r >= kTuneFirst && r < kTuneFirst + 4) {
            const int k = (r - kTuneFirst) % 2;  // 0 sorted, 1 unsorted
            tune.ms[0][k] += a;
            tune.rays[0][k] += host_cnt[kCntQueue];
            tune.ms[1][k] += b;
            tune.rays[1][k] += host_cnt[kCntSQueue];
            if (r == kTuneFirst + 3) {
                auto faster_sorted = [&](int q) {
                    return tune.rays[q][0] && tune.rays[q][1] &&
                           tune.ms[q][0] / double(tune.rays[q][0]) < tune.ms[q][1] / double(tune.rays[q][1]);
                };
                tune.sort_closest = faster_sorted(0);
                tune.sort_shadow = faster_sorted(1);
            }
        }
        if (r >= kOccFirst && r < kOccFirst + 6) {
            const int k = (r - kOccFirst) % 3;
            tune.occ_ms[k] += a + b;
            tune.occ_rays[k] += uint64_t(host_cnt[kCntQueue]) + host_cnt[kCntSQueue];
            if (r == kOccFirst + 5) {
                int best = 0;
                for (int q = 1; q < 3; ++q)
                    if (tune.occ_rays[q] && tune.occ_rays[best] &&
                        tune.occ_ms[q] / double(tune.occ_rays[q]) < tune.occ_ms[best] / double(tune.occ_rays[best]))
                        best = q;
                tune.quarters = kOccQuarters[best];
            }
        }
        rs.rays += host_cnt[kCntQueue];
        rs.shadow_rays += host_cnt[kCntSQueue];
        ++r;
        if (done) break;
        if (r > 100000) return hipErrorUnknown;
    }
    P.round = r;
    PT_CK(hipEventRecord(ev[1], stream));
    const uint32_t wh = S.width * S.height;
    if (gd && gd->ncoef) {
        reduce_shards_kernel<<<(gd->ncoef + kBlock - 1) / kBlock, kBlock, 0, stream>>>(gd->d_coef, gd->ncoef);
        PT_CK(hipGetLastError());
    }
    if (film_rgb) {
        launch_film(S, st.rec, st.rec_round, spp, film_rgb, film_aux, stream);
        PT_CK(hipGetLastError());
    }
    PT_CK(hipEventRecord(ev[2], stream));
    PT_CK(hipStreamSynchronize(stream));
    float tf = 0, tt = 0;
    hipEventElapsedTime(&tf, ev[1], ev[2]);
    hipEventElapsedTime(&tt, ev[0], ev[2]);
    rs.ms_total = tt;
    rs.ms_trace = t_trace;
    rs.ms_shadow = t_shadow;
    rs.ms_shade = t_shade;
    rs.ms_film = tf;
    rs.rounds = P.round;
    rs.sorted = (tune.sort_closest ? 1u : 0u) | (tune.sort_shadow ? 2u : 0u);
    rs.grid_quarters = tune.quarters;
    rs.paths = total;
    if (stats) *stats = rs;
    return hipSuccess;
#undef PT_CK
}

namespace {

// Every job of `jobs` in one shared pool (the batched wavefront).
hipError_t render_shared(const std::vector<BatchJob> &jobs, uint32_t pool, uint32_t rounds_per_launch, bool use_graph,
                         PathPool &pp, DevBuf &jobbuf, uint32_t *err, BatchStats *stats, hipStream_t stream,
                         uint32_t grid_quarters) {
    const uint32_t nj = uint32_t(jobs.size());
    if (nj == 0 || pool == 0 || rounds_per_launch == 0) return hipErrorInvalidValue;
    std::vector<JobDev> jd(nj);
    std::vector<WideTrace> jw(nj);
    std::vector<uint64_t> base(nj + 1);
    uint64_t total = 0;
    uint32_t spill_need = 0;
    for (uint32_t j = 0; j < nj; ++j) {
        const BatchJob &b = jobs[j];
        const uint32_t spd = uint32_t(sqrt(double(b.spp)));
        if (b.spp == 0 || spd * spd != b.spp || b.spp > kMaxSpp || b.vpool == 0) return hipErrorInvalidValue;
        if (b.max_depth + 3 > kRing) return hipErrorInvalidValue;
        const uint64_t n = uint64_t(b.S.width) * b.S.height * b.spp;
        jd[j].S = b.S;
        jd[j].base = total;
        jd[j].total = n;
        jd[j].spd = spd;
        jd[j].spp = b.spp;
        jd[j].max_depth = b.max_depth;
        jd[j].vpool = b.vpool;
        jw[j] = b.W;
        base[j] = total;
        total += n;
        spill_need = std::max(spill_need, b.W.spill_per_ray);
    }
    base[nj] = total;
    if (uint64_t(pool) > total) pool = uint32_t(total);
    if (total + pool >= 0xFFFFFFFFull) return hipErrorInvalidValue;  // the new-path counter runs up to total + pool
    hipError_t e;
#define PT_CK(x)                          \
    do {                                  \
        if ((e = (x)) != hipSuccess) return e; \
    } while (0)
    PT_CK(reserve_pool(pp, pool, total, true));
    PT_CK(reserve_ray_sort(pp.rsort, 2 * pool));  // the captured rounds allocate nothing
    WideTrace W{};
    W.spill_per_ray = spill_need;
    W.grid_quarters = grid_quarters;
    for (uint32_t j = 0; j < nj && !W.shapes; ++j) W.shapes = jw[j].shapes;  // the analytic-primitive kernels
    if (spill_need) {
        PT_CK(pp.spill.reserve(size_t(pool) * 2 * spill_need * sizeof(uint2)));
        W.spill = pp.spill.as<uint2>();
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t off_jw = align(nj * sizeof(JobDev)), off_base = off_jw + align(nj * sizeof(WideTrace));
    const size_t off_rounds = off_base + align((nj + 1) * sizeof(uint64_t));
    PT_CK(jobbuf.reserve(off_rounds + align(nj * 4)));
    char *jb = jobbuf.as<char>();
    const JobDev *d_jobs = reinterpret_cast<const JobDev *>(jb);
    const WideTrace *d_jw = reinterpret_cast<const WideTrace *>(jb + off_jw);
    const uint64_t *d_base = reinterpret_cast<const uint64_t *>(jb + off_base);
    uint32_t *d_rounds = reinterpret_cast<uint32_t *>(jb + off_rounds);
    PT_CK(hipMemcpyAsync(jb, jd.data(), nj * sizeof(JobDev), hipMemcpyHostToDevice, stream));
    PT_CK(hipMemcpyAsync(jb + off_jw, jw.data(), nj * sizeof(WideTrace), hipMemcpyHostToDevice, stream));
    PT_CK(hipMemcpyAsync(jb + off_base, base.data(), (nj + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    const State st = make_state(pp, err);
    Params P{total, 0u, 0u, 0u, 0u, pool, d_jobs, d_base, nj};
    const uint32_t grid = (pool + kBlock - 1) / kBlock;
    hipEvent_t ev[4];
    for (auto &x : ev) PT_CK(hipEventCreate(&x));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct Guard {
        hipEvent_t *e;
        hipGraph_t *g;
        hipGraphExec_t *x;
        ~Guard() {
            for (int k = 0; k < 4; ++k) hipEventDestroy(e[k]);
            if (*x) hipGraphExecDestroy(*x);
            if (*g) hipGraphDestroy(*g);
        }
    } guard{ev, &graph, &exec};
    PT_CK(hipMemsetAsync(st.counters, 0, 64, stream));
    const uint32_t first = pool;
    PT_CK(hipMemcpyAsync(st.counters + kCntNextId, &first, 4, hipMemcpyHostToDevice, stream));
    PT_CK(hipEventRecord(ev[0], stream));
    init_pool_multi_kernel<<<grid, kBlock, 0, stream>>>(st, P);
    PT_CK(hipGetLastError());
    auto enqueue_rounds = [&]() -> hipError_t {
        hipError_t e2;
        for (uint32_t k = 0; k < rounds_per_launch; ++k) {
            round_begin_kernel<<<1, 64, 0, stream>>>(st.counters);
            if ((e2 = hipGetLastError()) != hipSuccess) return e2;
            // both ray classes in (job, octant, origin Morton code) order
            uint32_t *q = nullptr;
            if ((e2 = sort_rays(W, reinterpret_cast<const float *>(st.rays), pool, pp.prim.as<uint32_t>(),
                                pp.thit.as<float>(), pp.bary.as<float>(), err, pp.rsort, st.counters + kCntQueue, &q,
                                stream, st.job, 0)) != hipSuccess)
                return e2;
            if ((e2 = trace_wide_multi_queued(W, d_jw, st.job, 0, reinterpret_cast<const float *>(st.rays), pool, 0,
                                              pp.prim.as<uint32_t>(), pp.thit.as<float>(), pp.bary.as<float>(), err,
                                              q, st.counters + kCntQueue, stream)) != hipSuccess)
                return e2;
            if ((e2 = sort_rays(W, reinterpret_cast<const float *>(st.srays), 2 * pool, pp.sprim.as<uint32_t>(),
                                nullptr, nullptr, err, pp.rsort, st.counters + kCntSQueue, &q, stream, st.job, 1)) !=
                hipSuccess)
                return e2;
            if ((e2 = trace_wide_multi_queued(W, d_jw, st.job, 1, reinterpret_cast<const float *>(st.srays),
                                              2 * pool, 1, pp.sprim.as<uint32_t>(), nullptr, nullptr, err, q,
                                              st.counters + kCntSQueue, stream)) != hipSuccess)
                return e2;
            shade_multi_kernel<<<grid, kBlock, 0, stream>>>(st, P);
            round_end_kernel<<<1, 64, 0, stream>>>(st.counters);
            if ((e2 = hipGetLastError()) != hipSuccess) return e2;
        }
        return hipSuccess;
    };
    if (use_graph) {
        PT_CK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const hipError_t ec = enqueue_rounds();
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(stream, &g);
        graph = g;
        PT_CK(ec);
        PT_CK(ee);
        PT_CK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    uint32_t host_cnt[16] = {};
    uint32_t launches = 0;
    // Every path ends within max_depth + 2 rounds of its start, so the pool drains in
    // at most (paths / pool + 1) * (max depth + 2) rounds.
    uint32_t dmax = 0;
    for (const auto &b : jobs) dmax = std::max(dmax, b.max_depth);
    const uint64_t round_cap = (total / pool + 2) * (uint64_t(dmax) + 3) + 16;
    for (;;) {
        if (use_graph) PT_CK(hipGraphLaunch(exec, stream));
        else PT_CK(enqueue_rounds());
        ++launches;
        PT_CK(hipMemcpyAsync(host_cnt, st.counters, 64, hipMemcpyDeviceToHost, stream));
        PT_CK(hipStreamSynchronize(stream));
        if (host_cnt[kCntDone]) break;
        if (uint64_t(launches) * rounds_per_launch > round_cap) return hipErrorUnknown;
    }
    PT_CK(hipEventRecord(ev[1], stream));
    schedule_kernel<<<nj, kSchedBlock, 0, stream>>>(d_jobs, st.rec_round, d_rounds, err);
    PT_CK(hipGetLastError());
    PT_CK(hipEventRecord(ev[2], stream));
    for (uint32_t j = 0; j < nj; ++j) {
        launch_film(jd[j].S, st.rec + jd[j].base, st.rec_round + jd[j].base, jd[j].spp, jobs[j].film_rgb,
                    jobs[j].film_aux, stream);
        PT_CK(hipGetLastError());
    }
    PT_CK(hipEventRecord(ev[3], stream));
    if (stats) {
        stats->job_rounds.assign(nj, 0);
        PT_CK(hipMemcpyAsync(stats->job_rounds.data(), d_rounds, nj * 4, hipMemcpyDeviceToHost, stream));
    }
    PT_CK(hipStreamSynchronize(stream));
    if (stats) {
        float a = 0, b = 0, c = 0, t = 0;
        hipEventElapsedTime(&a, ev[0], ev[1]);
        hipEventElapsedTime(&b, ev[1], ev[2]);
        hipEventElapsedTime(&c, ev[2], ev[3]);
        hipEventElapsedTime(&t, ev[0], ev[3]);
        stats->ms_rounds = a;
        stats->ms_schedule = b;
        stats->ms_film = c;
        stats->ms_total = t;
        stats->rounds = host_cnt[kCntRound];
        uint64_t acc[2];
        memcpy(acc, host_cnt + kCntAccRays, sizeof(acc));
        stats->rays = acc[0];
        stats->shadow_rays = acc[1];
        stats->paths = total;
        stats->launches = launches;
    }
    return hipSuccess;
#undef PT_CK
}

}  // namespace

hipError_t render_batch(const std::vector<BatchJob> &jobs, uint32_t pool, uint32_t rounds_per_launch, bool use_graph,
                        uint64_t solo_paths, PathPool &pp, DevBuf &jobbuf, uint32_t *err, BatchStats *stats,
                        hipStream_t stream, uint32_t grid_quarters) {
    const uint32_t nj = uint32_t(jobs.size());
    if (nj == 0) return hipErrorInvalidValue;
    hipEvent_t ev[2];
    hipError_t e;
    for (auto &x : ev)
        if ((e = hipEventCreate(&x)) != hipSuccess) return e;
    struct Guard {
        hipEvent_t *e;
        ~Guard() {
            hipEventDestroy(e[0]);
            hipEventDestroy(e[1]);
        }
    } guard{ev};
    BatchStats out;
    out.job_rounds.assign(nj, 0);
    std::vector<BatchJob> shared;
    std::vector<uint32_t> shared_idx;
    if ((e = hipEventRecord(ev[0], stream)) != hipSuccess) return e;
    for (uint32_t j = 0; j < nj; ++j) {
        const BatchJob &b = jobs[j];
        const uint64_t n = uint64_t(b.S.width) * b.S.height * b.spp;
        if (n >= solo_paths) {  // fills the pool by itself: the single-scene wavefront, in its own pool order
            RenderStats rs;
            e = render(b.S, b.W, b.spp, b.max_depth, b.vpool, pp, b.film_rgb, b.film_aux, err, &rs, stream);
            if (e != hipSuccess) return e;
            out.job_rounds[j] = uint32_t(rs.rounds);
            out.rounds += rs.rounds;
            out.rays += rs.rays;
            out.shadow_rays += rs.shadow_rays;
            out.paths += rs.paths;
            out.solo += 1;
        } else {
            shared.push_back(b);
            shared_idx.push_back(j);
        }
    }
    if (!shared.empty()) {
        BatchStats bs;
        e = render_shared(shared, pool, rounds_per_launch, use_graph, pp, jobbuf, err, &bs, stream, grid_quarters);
        if (e != hipSuccess) return e;
        for (size_t k = 0; k < shared.size(); ++k) out.job_rounds[shared_idx[k]] = bs.job_rounds[k];
        out.rounds += bs.rounds;
        out.rays += bs.rays;
        out.shadow_rays += bs.shadow_rays;
        out.paths += bs.paths;
        out.launches = bs.launches;
        out.ms_rounds = bs.ms_rounds;
        out.ms_schedule = bs.ms_schedule;
        out.ms_film = bs.ms_film;
    }
    if ((e = hipEventRecord(ev[1], stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    float t = 0;
    hipEventElapsedTime(&t, ev[0], ev[1]);
    out.ms_total = t;
    if (stats) *stats = out;
    return hipSuccess;
}

}  // namespace pt
}  // namespace pbrtm
