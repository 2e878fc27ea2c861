// Wavefront path tracer: device scene tables, path-pool state and the render
// driver. The pool holds `pool` path slots in structure-of-arrays form in HBM;
// each round traces every live slot's ray (closest hit) and the previous
// round's shadow rays (any hit) through the 8-wide BVH, then one shading kernel
// per slot resolves next-event estimation, applies the integrator's control
// logic (termination, emission with MIS), records finished paths for the film,
// starts new paths in freed slots and samples the BSDF and a light.
#pragma once

#include <string>
#include <vector>

#include "pbrtm_internal.h"
#include "pbrtm_shapes.h"
#include "wide_bvh.h"

namespace pbrtm {
namespace pt {

enum MatType : uint32_t { kDiffuse = 0, kConductor = 1, kDielectric = 2 };

struct DMat {  // 48 bytes
    uint32_t type;
    float c0, c1, c2;               // diffuse albedo sigmoid coefficients
    uint32_t eta_off, eta_n, k_off, k_n;  // conductor eta/k, or a dielectric's spectral eta (eta_n > 0)
    float ax, ay;                   // Trowbridge-Reitz alphas after the roughness remap
    float eta_c;                    // dielectric constant eta (when eta_n == 0)
    uint32_t pad;
};

constexpr uint32_t kInfiniteLight = 0xFFFFFFFFu;  // DLight::tri of a UniformInfiniteLight
constexpr uint32_t kMaxInfinite = 4;

struct DLight {  // 32 bytes: a DiffuseAreaLight on one triangle, or a UniformInfiniteLight
    uint32_t tri;     // kInfiniteLight: an infinite light (two_sided set; no "L": rgb_scale 1, c2 = inf)
    float scale;      // "scale" / to_photometric(L)
    float rgb_scale;  // RGBIlluminantSpectrum scale = 2 max(rgb)
    float c0, c1, c2; // sigmoid of rgb / rgb_scale
    uint32_t two_sided;
    uint32_t group;   // the scene's emitter (pbrtm_area_light index) the triangle belongs to
};

// Everything the kernels read about the scene; passed by value.
struct SceneDev {
    const float *tri;          // [n][9] render space, input order
    const uint2 *tri_info;     // x = material, y = light + 1 (0 = none) | flip << 31
    const DMat *mats;
    const DLight *lights;
    const float *light_pmf, *light_cdf;
    uint32_t num_tris, num_lights;
    const float *spec_lam, *spec_val;  // piecewise-linear spectrum pool
    uint32_t cie_off[3], cie_n, d65_off, d65_n;
    // uniform tails of the CIE and D65 wavelength grids (start index, 1 / step) for the
    // interval guess; cie_shared: x, y and z use the same grid
    uint32_t cie_uni, d65_uni, cie_shared;
    float cie_inv_step, d65_inv_step;
    float cam_from_raster[16], render_from_camera[16];
    float cam_origin[3];
    const float *filt_f;         // [64][64] filter values
    const float *filt_cond_func; // [64][64] |f|
    const float *filt_cond_cdf;  // [64][65]
    const float *filt_cond_int;  // [64]
    const float *filt_marg_func; // [64]
    const float *filt_marg_cdf;  // [65]
    float filt_marg_int, filt_radius;
    uint32_t filt_n;
    float out_from_sensor[9];
    uint32_t width, height;
    float ao_scale;  // ambient occlusion: 1 / the colour space illuminant's to_photometric(Y)
    float sn_c[9];   // surface normal: sigmoid coefficients of RGBAlbedoSpectrum(0.01 e_c), c = r, g, b
    uint32_t num_inf, inf_idx[kMaxInfinite];  // the infinite lights' indices into lights, in light order
    float inf_dist;  // 2 x the scene's bounding-sphere radius (UniformInfiniteLight::sample_li)
    const DShape *shapes;        // analytic primitives (spheres, disks, instanced triangles); null if none
    const uint32_t *prim_shape;  // per primitive: index into shapes + 1, 0 = a triangle; null if none
};

// Integrators render() runs (SceneBuilder::build_integrator, scene_builder.cu:289-313).
enum Integrator : uint32_t { kIntPath = 0, kIntMegakernelPath = 1, kIntAmbientOcclusion = 2, kIntSurfaceNormal = 3 };

// Gradient replay state (device pointers).
struct GradDev {
    const float4 *lfin, *pfin;  // per path id: final L and wavelength pdfs of the forward pass
    const float *adj;           // per pixel [3]: dLoss/d(rgb_sum / weight_sum) through the output matrix
    float *d_coef;              // kGradShards x [materials][3]: dLoss/d(sigmoid coefficients); the
                                // render leaves the sum in the first copy
    uint32_t ncoef;             // 3 x materials
    double *acc;                // [ngroups + 1]: scale-linear part per emitter, then the pmf part
    float4 *g, *ltot, *nee;     // per slot: dLoss/dL, the final L, pending NEE (G, w, group)
    uint32_t ngroups;
};

constexpr uint32_t kGradShards = 32;

// One job of a batched render, as the kernels read it.
struct JobDev {
    SceneDev S;
    uint64_t base, total;  // first global path id, paths (W * H * spp)
    uint32_t spd, spp, max_depth, vpool;  // vpool: the one-by-one pool size whose schedule the film reproduces
};

// Host-side scene set-up (render space, spectra, camera, filter, lights).
struct SceneHost {
    std::vector<float> tri;  // render space
    std::vector<uint2> tri_info;
    std::vector<DMat> mats;
    std::vector<DLight> lights;
    std::vector<float> light_pmf, light_cdf;
    std::vector<float> spec_lam, spec_val;
    uint32_t cie_off[3], cie_n, d65_off, d65_n;
    uint32_t cie_uni, d65_uni, cie_shared;
    float cie_inv_step, d65_inv_step;
    float cam_from_raster[16], render_from_camera[16], cam_origin[3];
    std::vector<float> filt_f, filt_cond_func, filt_cond_cdf, filt_cond_int, filt_marg_func, filt_marg_cdf;
    float filt_marg_int, filt_radius;
    uint32_t filt_n;
    float out_from_sensor[9];
    uint32_t width, height;
    float ao_scale;
    float sn_c[9];
    uint32_t num_inf, inf_idx[kMaxInfinite];
    float inf_dist;
    std::vector<DShape> shapes;      // analytic primitives
    std::vector<uint32_t> prim_shape;  // per primitive: shape index + 1, 0 = triangle (empty: no shapes)
    std::vector<float> mat_rgb;      // per material [3]: the diffuse albedo RGB (gradients)
    std::vector<float> group_scale;  // per emitter: the scene's "scale" parameter (gradients)
};

struct SceneDesc;  // the C-ABI description, see include/pbrtm.h

// Builds the host tables; returns an empty string or an error message.
std::string build_scene_host(const void *desc, SceneHost &out);

// RGB-to-spectrum sigmoid coefficients of an albedo RGB, and the sRGB matrices (for tests).
void rgb_to_sigmoid(const float rgb[3], float out[3]);
// d(c0, c1, c2) / d(r, g, b) of rgb_to_sigmoid, row-major [q][k].
void rgb_to_sigmoid_jacobian(const float rgb[3], double J[9]);
// out[m] = J(rgb[m])^T coef[m] for n RGBs (multi-threaded).
void rgb_to_sigmoid_jacobian_transpose(const float *rgb, const float *coef, size_t n, float *out);
void srgb_matrices(float xyz_from_rgb[9], float rgb_from_xyz[9], float white_xy[2]);
// SquareMatrix<4>::inverse (euclidean_space/squared_matrix.cu), row-major; false if singular.
bool mat4_inverse_ref(const float in[16], float out[16]);

struct PathPool {
    DevBuf path_id, flags, rng, dim, lam, lam_pdf, L, beta, nee_L, scal, prev, rays, prim, thit, bary, srays, sprim,
        rec, rec_round, counters, queue, squeue, spill, job, sround;
    RaySortScratch rsort;
    uint32_t cap = 0;
    uint64_t rec_cap = 0;
};

struct RenderStats {
    double ms_total = 0, ms_trace = 0, ms_shadow = 0, ms_shade = 0, ms_film = 0;
    uint64_t rounds = 0, rays = 0, shadow_rays = 0, paths = 0;
    uint32_t sorted = 0;  // the ray-order autotuner's choice: 1 closest-hit rays sorted, 2 shadow rays sorted
    uint32_t grid_quarters = 4;  // the occupancy autotuner's choice: traversal grid in quarters of the maximum
};

// Renders the whole image; film_rgb (device, [h][w][3]) and film_aux (device,
// [h][w][4], nullable) receive RGBFilm::get_pixel_rgb and (rgb_sum, weight_sum).
// lfin / pfin (nullable, per path id) receive each path's final L and wavelength
// pdfs; with gd the render is a gradient replay (film_rgb may then be null).
// counts (device, nullable, 6 words zeroed by the caller): the traversal's totals,
// [visits, box tests, primitive tests] of the closest-hit then the shadow rays,
// from the counting instance of the traversal kernel.
hipError_t render(const SceneDev &S, const WideTrace &W, uint32_t spp, uint32_t max_depth, uint32_t pool,
                  PathPool &pp, float *film_rgb, float *film_aux, uint32_t *err, RenderStats *stats,
                  hipStream_t stream, float4 *lfin = nullptr, float4 *pfin = nullptr,
                  const GradDev *gd = nullptr, uint32_t integrator = kIntPath,
                  unsigned long long *counts = nullptr);

// Marks the leaf positions of analytic primitives for the traversal: the word after
// a position's nine coordinates receives prim_shape[prim_order[k]] (shape index + 1).
hipError_t tag_leaf_shapes(float *tri_leaf, const uint32_t *prim_order, const uint32_t *prim_shape, uint32_t n,
                           hipStream_t stream);

// Loss sum((film - target)^2) per pixel into loss_px, and adj [h][w][3] =
// dLoss/d(rgb_sum / weight_sum) through out_from_sensor (divided by the
// weight sum where the film divides by it).
hipError_t film_adjoint(const SceneDev &S, const float *film_rgb, const float *film_aux, const float *target,
                        float *adj, double *loss, hipStream_t stream);

struct BatchJob {
    SceneDev S;
    WideTrace W;
    uint32_t spp, max_depth, vpool;
    float *film_rgb, *film_aux;  // device; film_aux nullable
};

struct BatchStats {
    double ms_total = 0, ms_rounds = 0, ms_schedule = 0, ms_film = 0;
    uint64_t rounds = 0, rays = 0, shadow_rays = 0, paths = 0;
    uint32_t launches = 0, solo = 0;
    std::vector<uint32_t> job_rounds;  // each job's round count in its one-by-one schedule
};

// Renders the jobs with at least solo_paths paths one after another with the
// single-scene wavefront (such a job fills the GPU by itself, and mixing it with
// others only shares the caches between trees), then every other job in one
// shared pool of `pool` slots; `rounds_per_launch` rounds are enqueued (as one
// captured graph when use_graph) between checks of the device's done flag.
// Each job's film equals its one-by-one render with pool size vpool.
// grid_quarters: the shared pool's persistent traversal grids, in quarters of the
// resident maximum (0 = all).
hipError_t render_batch(const std::vector<BatchJob> &jobs, uint32_t pool, uint32_t rounds_per_launch, bool use_graph,
                        uint64_t solo_paths, PathPool &pp, DevBuf &jobbuf, uint32_t *err, BatchStats *stats,
                        hipStream_t stream, uint32_t grid_quarters = 0);

}  // namespace pt
}  // namespace pbrtm
