// HLBVH build: device data structures and the host entry point.
#pragma once

#include "pbrtm_device.h"
#include "pbrtm_internal.h"

namespace pbrtm {

enum CounterSlot : int {
    kCntSmallTop = 0,
    kCntSmallBot = 1,
    kCntLargeTop0 = 2,
    kCntLargeTop1 = 3,
    kCntLargeBot0 = 4,
    kCntLargeBot1 = 5,
    kCntTasks = 6,
    kCntTreelets = 7,
    kCntErr = 8,
    kCntLevelsTop = 9,
    kCntLevelsBot = 10,
    kCntTopItems0 = 11,  // fused top level: nodes and chunk tasks of the two level parities
    kCntTopItems1 = 12,
    kCntTopTasks0 = 13,
    kCntTopTasks1 = 14,
    kCntLaneBot0 = 15,   // treelets of 2..8 and 9..16 primitives for the one-lane builder
    kCntLaneBot1 = 16,
    kCntMidTop = 17,     // top-level nodes of 65..2048 treelets for the workgroup finisher
    kNumCounters = 18,
};

// Everything the build kernels write, passed by value as a kernel argument.
struct BuildArgs {
    BNode *nodes;
    uint32_t *prim_order;
    Rec *trec[2];
    Rec *prec[2];
    WorkItem *small_top, *small_bot;
    WorkItem *large_top[2], *large_bot[2];
    WorkItem *lane_bot[2];
    WorkItem *mid_top;
    uint32_t cap_small_top, cap_small_bot, cap_large_top, cap_large_bot;
    uint32_t cap_lane_bot[2];
    uint32_t cap_mid_top;
    uint32_t *counters;
    uint32_t *err;
};

// Scratch reused across builds on one context (grown, never shrunk).
struct BuildScratch {
    DevBuf counters, hist, keys[2], vals[2], prec[2], trec[2];
    DevBuf lists, acc, splits, tasks, task_off, task_left, task_lprims;
    DevBuf top_items, top_acc, top_tasks, top_hist, top_left, top_bkt;  // fused top level
    DevBuf scan_scratch;
    DevBuf tri_sorted;  // the triangles in sorted (Morton) order, 9 floats each
    SortScratch sort;
    uint32_t *host_counters = nullptr;
    const uint32_t *sorted_pid = nullptr;  // input index per sorted position (one of vals[])
    void release();
};

// The built tree: binary nodes in preorder-reserved slots (2N-1 of them, root 0)
// and the primitive order that leaves index into.
struct BvhDevice {
    uint32_t n = 0;
    uint32_t num_treelets = 0;
    uint32_t top_levels = 0, bottom_levels = 0;
    uint32_t small_top = 0, small_bottom = 0;
    DevBuf nodes, prim_order;
};

struct BuildPhaseTimes {
    hipEvent_t ev[6];  // start, sorted, treelets, top done, bottom done, end
};

// Centroid bounds, Morton codes and the stable sort; *keys_out/*vals_out name
// the scratch buffers holding sorted codes and input indices.
hipError_t morton_sort_device(const float *d_tri, uint32_t n, BuildScratch &S, uint32_t **keys_out,
                              uint32_t **vals_out, hipStream_t stream);

hipError_t build_hlbvh_device(const float *d_tri, uint32_t n, BuildScratch &S, BvhDevice &out,
                              uint32_t *device_error, hipStream_t stream, BuildPhaseTimes *times);

// After build_hlbvh_device on the same scratch and stream: the primitive order
// becomes input indices (the build leaves sorted positions in it), and tri_leaf
// ([n][12] floats) receives the triangles in leaf order.
hipError_t leaf_order(const BuildScratch &S, BvhDevice &bvh, float *tri_leaf, hipStream_t stream);

}  // namespace pbrtm
