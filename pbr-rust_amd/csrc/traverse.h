// Traversal-side device views and launchers.
#pragma once

#include "pbrtm_device.h"
#include "pbrtm_internal.h"

namespace pbrtm {

// Everything a traversal kernel reads: binary nodes, leaf-ordered triangles
// (9 floats in a 12-float slot per leaf position) and the leaf position -> primitive id map.
struct BvhTrace {
    const BNode *nodes;
    const float *tri_leaf;
    const uint32_t *prim_order;
    uint32_t num_prims;
};

// rays: [n][8] floats (o.xyz, t_max, d.xyz, unused); t_max <= 0 marks an inactive ray.
// counts (device, nullable): adds [nodes visited, box tests, triangle tests] of this
// launch; in the binary traversal a visit is one box test.
hipError_t trace_binary(const BvhTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                        float *t_out, float *bary_out, uint32_t *err, unsigned long long *counts,
                        hipStream_t stream);
hipError_t deindex_triangles(const float *verts, const uint32_t *idx, uint32_t nt, uint32_t nv, float *tri,
                             uint32_t *err, hipStream_t stream);
hipError_t ao_bounce(const float *tri, uint32_t num_tri, const float *rays, const uint32_t *prim, const float *bary,
                     uint32_t n, float *rays_out, hipStream_t stream);

}  // namespace pbrtm
