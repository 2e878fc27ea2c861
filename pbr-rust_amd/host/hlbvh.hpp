// C++ host interface mirroring the reference's aggregate (accelerator/hlbvh.h:76-95):
// HLBVH::create, bounds, intersect, fast_intersect — batched, because on the
// GPU one call serves a whole wavefront of rays. Header-only over the C ABI in
// include/pbrtm.h; link with -lpbrtm.
//
// Error behaviour: the reference aborts the process (gpu/macro.h:19-51); here
// failures throw std::runtime_error with the library's message, as the
// reference's scene front end does for bad input (scene_builder.cu:499).
#pragma once

#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbrtm.h"

namespace pbrtm_host {

struct Point3f {
    float x, y, z;
};

struct Bounds3f {  // bounds3.h: empty = (+max, lowest)
    Point3f p_min, p_max;
};

struct Ray {  // base/ray.h:8-19, plus the t_max the reference passes alongside
    Point3f o;
    Point3f d;
    float t_max = std::numeric_limits<float>::infinity();
};

// ShapeIntersection's t_hit plus the hit primitive and barycentrics
// (shapes/triangle.h:15-22); prim == PBRTM_NO_HIT means "no intersection".
struct Hit {
    uint32_t prim;
    float t_hit;
    float b0, b1, b2;
};

class Context {
  public:
    explicit Context(int device = 0) {
        const int rc = pbrtm_create(&ctx_, device);
        if (rc != PBRTM_OK) throw std::runtime_error("pbrtm_create failed: " + std::to_string(rc));
    }
    ~Context() { pbrtm_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    pbrtm_ctx *get() const { return ctx_; }
    void check(int rc) const {
        if (rc != PBRTM_OK) throw std::runtime_error(std::string("pbrtm: ") + pbrtm_last_error(ctx_));
    }

  private:
    pbrtm_ctx *ctx_ = nullptr;
};

class HLBVH {
  public:
    // HLBVH::create(primitives, allocator) (hlbvh.h:76-77): a TriangleMesh's render-space
    // points and vertex indices (triangle_mesh.cu:39-106); the aggregate owns device copies.
    static std::unique_ptr<HLBVH> create(Context &ctx, const std::vector<Point3f> &points,
                                         const std::vector<uint32_t> &vertex_indices,
                                         pbrtm_build_stats *stats = nullptr) {
        if (vertex_indices.size() % 3 != 0) throw std::runtime_error("vertex_indices: not a triangle list");
        pbrtm_bvh *h = nullptr;
        ctx.check(pbrtm_bvh_build(ctx.get(), reinterpret_cast<const float *>(points.data()), points.size(),
                                  vertex_indices.data(), vertex_indices.size() / 3, 0, &h, stats));
        return std::unique_ptr<HLBVH>(new HLBVH(ctx, h));
    }

    ~HLBVH() { pbrtm_bvh_free(bvh_); }
    HLBVH(const HLBVH &) = delete;
    HLBVH &operator=(const HLBVH &) = delete;

    // HLBVH::bounds() (hlbvh.h:79-86).
    Bounds3f bounds() const {
        float b[6];
        ctx_.check(pbrtm_bvh_bounds(bvh_, b));
        return {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
    }

    // HLBVH::intersect(ray, t_max) for every ray (hlbvh.cu:231-289).
    std::vector<Hit> intersect(const std::vector<Ray> &rays) const {
        std::vector<float> packed = pack(rays);
        const size_t n = rays.size();
        std::vector<uint32_t> prim(n);
        std::vector<float> t(n), bary(3 * n);
        ctx_.check(pbrtm_bvh_intersect(bvh_, packed.data(), n, 0, 0, prim.data(), t.data(), bary.data(), nullptr));
        std::vector<Hit> out(n);
        for (size_t i = 0; i < n; ++i) out[i] = {prim[i], t[i], bary[3 * i], bary[3 * i + 1], bary[3 * i + 2]};
        return out;
    }

    // HLBVH::fast_intersect(ray, t_max) for every ray (hlbvh.cu:179-229).
    std::vector<bool> fast_intersect(const std::vector<Ray> &rays) const {
        std::vector<float> packed = pack(rays);
        const size_t n = rays.size();
        std::vector<uint32_t> prim(n);
        ctx_.check(pbrtm_bvh_intersect(bvh_, packed.data(), n, 1, 0, prim.data(), nullptr, nullptr, nullptr));
        std::vector<bool> out(n);
        for (size_t i = 0; i < n; ++i) out[i] = prim[i] != PBRTM_NO_HIT;
        return out;
    }

    pbrtm_bvh *handle() const { return bvh_; }

  private:
    HLBVH(Context &ctx, pbrtm_bvh *h) : ctx_(ctx), bvh_(h) {}

    static std::vector<float> pack(const std::vector<Ray> &rays) {
        std::vector<float> p(8 * rays.size());
        for (size_t i = 0; i < rays.size(); ++i) {
            const Ray &r = rays[i];
            float *q = &p[8 * i];
            q[0] = r.o.x; q[1] = r.o.y; q[2] = r.o.z; q[3] = r.t_max;
            q[4] = r.d.x; q[5] = r.d.y; q[6] = r.d.z; q[7] = 0.f;
        }
        return p;
    }

    Context &ctx_;
    pbrtm_bvh *bvh_;
};

}  // namespace pbrtm_host
