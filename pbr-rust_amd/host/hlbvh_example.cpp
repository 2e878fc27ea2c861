// Builds the aggregate of a unit cube through the C++ interface, shoots six
// axis-aligned rays at its faces and prints "prim t" per ray (a miss prints -1).
#include <cstdio>

#include "hlbvh.hpp"

int main() {
    using namespace pbrtm_host;
    try {
        Context ctx(0);
        std::vector<Point3f> p;
        for (int i = 0; i < 8; ++i) p.push_back({float(i & 1), float((i >> 1) & 1), float((i >> 2) & 1)});
        const std::vector<uint32_t> idx = {0, 1, 3, 0, 3, 2, 4, 7, 5, 4, 6, 7, 0, 4, 5, 0, 5, 1,
                                           2, 3, 7, 2, 7, 6, 0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3};
        auto bvh = HLBVH::create(ctx, p, idx);
        const Bounds3f b = bvh->bounds();
        std::printf("bounds %g %g %g %g %g %g\n", b.p_min.x, b.p_min.y, b.p_min.z, b.p_max.x, b.p_max.y, b.p_max.z);
        std::vector<Ray> rays = {
            {{0.25f, 0.5f, -1.f}, {0, 0, 1}}, {{0.25f, 0.5f, 2.f}, {0, 0, -1}}, {{-1.f, 0.5f, 0.25f}, {1, 0, 0}},
            {{2.f, 0.5f, 0.25f}, {-1, 0, 0}}, {{0.5f, -1.f, 0.25f}, {0, 1, 0}}, {{3.f, 3.f, 3.f}, {1, 0, 0}},
        };
        const std::vector<Hit> hits = bvh->intersect(rays);
        const std::vector<bool> any = bvh->fast_intersect(rays);
        for (size_t i = 0; i < hits.size(); ++i)
            std::printf("%d %g %d\n", hits[i].prim == PBRTM_NO_HIT ? -1 : int(hits[i].prim), hits[i].t_hit,
                        int(any[i]));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
