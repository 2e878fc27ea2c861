// 8-wide BVH collapsed from the binary build tree, and its traversal.
#pragma once

#include "pbrtm_device.h"
#include "pbrtm_internal.h"

namespace pbrtm {

namespace pt {
struct DShape;
}

constexpr int kWide = 8;
constexpr uint32_t kWideLeaf = 0x80000000u;
constexpr uint32_t kWideEmpty = 0xFFFFFFFFu;

// 128 bytes = one cache line. Child boxes are quantized to 8 bits per plane
// against the node's origin (the union's lower corner) and a power-of-two step
// per axis: plane = fmaf(q, 2^exp, org), rounded outwards at encoding so that
// every decoded box contains the child's exact FP32 box. Then one (word0,
// word1) reference per child slot. Interior child: word0 = wide node index,
// word1 = 0. Leaf child: word0 = kWideLeaf | first leaf position, word1 =
// primitive count. Unused slot: word0 = kWideEmpty, inverted box. Slots hold
// the collapsed binary subtree's frontier in left-to-right order.
struct alignas(128) WNode {
    float org[3];
    int8_t exp[3];
    uint8_t pad;
    uint8_t q[6][kWide];  // lo_x, lo_y, lo_z, hi_x, hi_y, hi_z per slot
    uint32_t ref[kWide][2];
};
static_assert(sizeof(WNode) == 128, "WNode is one cache line");

// Per wide node and ray octant (bit k = direction k negative): the order in
// which the reference's depth-first, near-child-first binary traversal reaches
// the slots, as 8 nibbles (position p -> slot), 0xF past the last slot.
struct WideBvh {
    DevBuf nodes, order;
    uint32_t num_nodes = 0;
    uint32_t max_depth = 0;   // wide levels on the longest root-to-leaf path
    uint32_t collapse_levels = 0;
};

// Collapses the binary tree (preorder-reserved ids, root 0, n primitives).
hipError_t collapse_wide(const BNode *bnodes, uint32_t n, WideBvh &out, DevBuf &scratch, uint32_t *err,
                         hipStream_t stream);

struct WideTrace {
    const WNode *nodes;
    const uint32_t *order;
    const float *tri_leaf;
    const uint32_t *prim_order;
    const pt::DShape *shapes;  // analytic primitives (leaf pad word = index + 1), null for triangle-only trees
    uint32_t num_nodes, num_prims, max_depth;
    uint2 *spill;           // per-ray stack overflow area (max_depth - kLdsStack entries per ray), may be null
    uint32_t spill_per_ray;
    float lo[3], hi[3];     // bounds of the whole tree (ray ordering)
    uint32_t grid_quarters; // persistent traversal grid: this many quarters of the resident maximum (0 = all)
};

// counts (device, nullable): adds [wide nodes visited, child box tests, triangle tests].
// queue (device, n entries) and queue_len (device, zeroed by the caller) hold the
// compacted active rays.
hipError_t trace_wide(const WideTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                      float *t_out, float *bary_out, uint32_t *err, unsigned long long *counts, uint32_t *queue,
                      uint32_t *queue_len, hipStream_t stream);

// As trace_wide, for rays of several jobs: ray i is traced against jobs[ray_job[i >> job_shift]]
// (nodes, order, tri_leaf, prim_order, num_nodes, num_prims); B supplies the spill area.
hipError_t trace_wide_multi(const WideTrace &B, const WideTrace *jobs, const uint32_t *ray_job, uint32_t job_shift,
                            const float *rays, uint32_t n, int any_hit, uint32_t *prim_out, float *t_out,
                            float *bary_out, uint32_t *err, uint32_t *queue, uint32_t *queue_len,
                            hipStream_t stream);

// Orders the active rays of [n] by (direction octant, Morton code of the origin
// on a 64^3 grid over the tree's box; 20-bit keys, two radix passes) into *queue
// (a pointer into the scratch, n words) and counts them into *queue_len (zeroed
// by the caller); inactive rays report a miss. trace_wide_queued then traces that
// queue.
struct RaySortScratch {
    DevBuf keys[2], vals[2], hist;
    SortScratch sort;
};
// ray_job (nullable): rays of a batch, ray i of job ray_job[i >> job_shift]; the
// job then leads the key (three radix passes). Every step is a kernel (no memset),
// so that a graph may capture the call once reserve_ray_sort has sized the scratch.
hipError_t sort_rays(const WideTrace &B, const float *rays, uint32_t n, uint32_t *prim_out, float *t_out,
                     float *bary_out, uint32_t *err, RaySortScratch &s, uint32_t *queue_len, uint32_t **queue,
                     hipStream_t stream, const uint32_t *ray_job = nullptr, uint32_t job_shift = 0);
hipError_t reserve_ray_sort(RaySortScratch &s, uint32_t n);
hipError_t trace_wide_multi_queued(const WideTrace &B, const WideTrace *jobs, const uint32_t *ray_job,
                                   uint32_t job_shift, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                                   float *t_out, float *bary_out, uint32_t *err, const uint32_t *queue,
                                   const uint32_t *queue_len, hipStream_t stream);
// counts (device, nullable): adds [nodes visited, child box tests, primitive tests].
hipError_t trace_wide_queued(const WideTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                             float *t_out, float *bary_out, uint32_t *err, const uint32_t *queue,
                             const uint32_t *queue_len, hipStream_t stream, unsigned long long *counts = nullptr);

constexpr uint32_t kLdsStack = 12;

}  // namespace pbrtm
