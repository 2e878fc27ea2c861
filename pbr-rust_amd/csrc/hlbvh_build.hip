// HLBVH construction on gfx950, reproducing the tree of accelerator/hlbvh.cu
// (reference): 30-bit Morton codes over cube-adjusted centroid bounds
// (hlbvh.cu:127-141, :428-446), treelets = runs of equal code >> 9
// (hlbvh.cu:36-42, :504-543), a 24-bucket SAH top level over treelets
// (hlbvh.cu:649-826) and a midpoint bottom level inside each treelet
// (hlbvh.cu:291-382, :828-857).
//
// Layout and scheduling are MI355X-native, not the reference's:
//  * one fused pass computes the Morton code and all three radix histograms;
//  * a stable onesweep radix sort orders (code, index), which refines the
//    reference's treelet-only counting sort into a canonical order;
//  * every split, top or bottom, runs on 32-byte records in two ping-pong
//    buffers. Nodes with more than 64 elements split level by level with one
//    workgroup per 2048-element chunk (atomic min/max of order-mapped floats,
//    chunked stable partition); nodes with at most 64 elements finish their
//    whole subtree inside one wave (LDS records, ballot partition);
//  * node ids are preassigned in preorder from primitive counts (a node with P
//    primitives owns 2P-1 ids: left = id + 1, right = id + 2*P_left), so the
//    output does not depend on the order in which work items run.
#include "hlbvh_build.h"
#include "wave_ops.h"

namespace pbrtm {

namespace {

constexpr int kChunk = 2048;       // elements per workgroup task
constexpr uint32_t kMidNode = kChunk;  // top-level nodes up to this size finish in one workgroup
constexpr int kChunkBlock = 256;
constexpr int kChunkItems = kChunk / kChunkBlock;
constexpr int kAccWords = 192;     // per-item accumulator (see Acc below)

// Per large work item: centroid box, full box, then 24 buckets of (count, box),
// floats stored order-mapped (f2ord) so integer atomics give exact min/max.
struct Acc {
    uint32_t cb[6];
    uint32_t fb[6];
    uint32_t bcnt[kNumBuckets];
    uint32_t bbox[kNumBuckets][6];
    uint32_t pad[kAccWords - 12 - kNumBuckets * 7];
};
static_assert(sizeof(Acc) == kAccWords * 4, "Acc layout");

struct Split {
    uint32_t axis;
    uint32_t bucket;     // top: last bucket on the left side
    float split_val;     // midpoint predicate value (bottom, and top fallback)
    uint32_t mode;       // 0 = SAH bucket predicate, 1 = midpoint predicate
    float base, span;    // top: bucket mapping
    uint32_t left_prims; // top: primitives on the left side
    uint32_t pad;
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ Rec *pick(Rec *const (&pair)[2], uint32_t b) { return b ? pair[1] : pair[0]; }

// The two ping-pong record buffers of one level kind, passed by value as a kernel
// argument (the host never reads device memory to find them).
struct RecPair {
    Rec *buf[2];
};

__device__ __forceinline__ Box acc_box(const uint32_t *w) {
    return {{ord2f(w[0]), ord2f(w[1]), ord2f(w[2])}, {ord2f(w[3]), ord2f(w[4]), ord2f(w[5])}};
}

// 24 * ((c - base) / span), truncated, 24 folded to 23 (hlbvh.cu:693-701).
__device__ __forceinline__ uint32_t bucket_of(float c, float base, float span) {
    uint32_t b = f2u_sat(float(kNumBuckets) * ((c - base) / span));
    return b >= kNumBuckets ? kNumBuckets - 1 : b;
}

__device__ __forceinline__ Rec make_rec(Box b, uint32_t a, uint32_t cnt) {
    Rec r;
    r.lo[0] = b.lo.x; r.lo[1] = b.lo.y; r.lo[2] = b.lo.z;
    r.hi[0] = b.hi.x; r.hi[1] = b.hi.y; r.hi[2] = b.hi.z;
    r.a = a;
    r.b = cnt;
    return r;
}

__device__ __forceinline__ void write_leaf(BNode *nodes, uint32_t id, Box b, uint32_t first, uint32_t count) {
    BNode nd;
    nd.lo[0] = b.lo.x; nd.lo[1] = b.lo.y; nd.lo[2] = b.lo.z;
    nd.hi[0] = b.hi.x; nd.hi[1] = b.hi.y; nd.hi[2] = b.hi.z;
    nd.a = kLeafFlag | first;
    nd.b = count;
    nodes[id] = nd;
}

__device__ __forceinline__ void write_interior(BNode *nodes, uint32_t id, Box b, uint32_t left_prims,
                                               uint32_t axis) {
    BNode nd;
    nd.lo[0] = b.lo.x; nd.lo[1] = b.lo.y; nd.lo[2] = b.lo.z;
    nd.hi[0] = b.hi.x; nd.hi[1] = b.hi.y; nd.hi[2] = b.hi.z;
    nd.a = id + 1;
    nd.b = (id + 2 * left_prims) | (axis << 30);
    nodes[id] = nd;
}

__device__ __forceinline__ void push_item(WorkItem *list, uint32_t *counter, uint32_t cap, WorkItem it,
                                          uint32_t *err) {
    uint32_t slot = atomicAdd(counter, 1u);
    if (slot < cap) list[slot] = it;
    else atomicOr(err, uint32_t(kErrCapacity));
}

// Wave-aggregated append: one atomic per wave and list instead of one per item.
// Every active lane of the wave must call it (it ballots).
__device__ __forceinline__ void wave_push(bool want, const WorkItem &it, WorkItem *list, uint32_t *counter,
                                          uint32_t cap, uint32_t *err) {
    const uint64_t m = __ballot(want);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(m)));
    base = __shfl(base, leader, kWave);
    if (want) {
        const uint32_t slot = base + __popcll(m & lanemask_lt());
        if (slot < cap) list[slot] = it;
        else atomicOr(err, uint32_t(kErrCapacity));
    }
}

// A treelet that became a top-level leaf is the root of a bottom-level build
// over its primitive range (hlbvh.cu:565-607 expands such leaves).
__device__ void emit_bottom_root(const BuildArgs &A, const Rec &treelet, uint32_t node) {
    const uint32_t first = treelet.a, count = treelet.b;
    if (count == 1) {
        const Rec p = A.prec[0][first];
        write_leaf(A.nodes, node, rec_box(p), first, 1);
        A.prim_order[first] = p.a;
        return;
    }
    WorkItem it{first, first + count, node, count, 0, {0, 0, 0}};
    if (count <= kSmallNode) push_item(A.small_bot, A.counters + kCntSmallBot, A.cap_small_bot, it, A.err);
    else push_item(A.large_bot[0], A.counters + kCntLargeBot0, A.cap_large_bot, it, A.err);
}

// ---------------------------------------------------------------- Morton pass

__global__ __launch_bounds__(256) void centroid_bounds_kernel(const float *__restrict__ tri, uint32_t n,
                                                              uint32_t *__restrict__ cbounds,
                                                              uint32_t *__restrict__ err) {
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    bool bad = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *t = tri + 9ull * i;
        f3 p0{t[0], t[1], t[2]}, p1{t[3], t[4], t[5]}, p2{t[6], t[7], t[8]};
#pragma unroll
        for (int k = 0; k < 9; ++k) bad |= !isfinite(t[k]);
        f3 c = box_centroid(tri_box(p0, p1, p2));
        lo[0] = min(lo[0], f2ord(c.x)); lo[1] = min(lo[1], f2ord(c.y)); lo[2] = min(lo[2], f2ord(c.z));
        hi[0] = max(hi[0], f2ord(c.x)); hi[1] = max(hi[1], f2ord(c.y)); hi[2] = max(hi[2], f2ord(c.z));
    }
    __shared__ uint32_t s[4][6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min_u32(lo[k]);
        hi[k] = wave_max_u32(hi[k]);
    }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        for (int k = 0; k < 3; ++k) { s[w][k] = lo[k]; s[w][3 + k] = hi[k]; }
    }
    if (__ballot(bad) && __lane_id() == 0) atomicOr(err, uint32_t(kErrNonFinite));
    __syncthreads();
    if (threadIdx.x < 6) {
        uint32_t v = s[0][threadIdx.x];
        for (int ww = 1; ww < 4; ++ww)
            v = threadIdx.x < 3 ? min(v, s[ww][threadIdx.x]) : max(v, s[ww][threadIdx.x]);
        if (threadIdx.x < 3) atomicMin(&cbounds[threadIdx.x], v);
        else atomicMax(&cbounds[threadIdx.x], v);
    }
}

// Cube-adjusted centroid bounds (hlbvh.cu:428-446), identical in every thread.
__device__ __forceinline__ Box morton_frame(const uint32_t *cbounds) {
    Box cb = acc_box(cbounds);
    const int md = box_max_dim(cb);
    const float radius = (sel3(cb.hi, md) - sel3(cb.lo, md)) / 2;
    const f3 c = box_centroid(cb);
    f3 lo = cb.lo, hi = cb.hi;
    if (md != 0) { lo.x = c.x - radius; hi.x = c.x + radius; }
    if (md != 1) { lo.y = c.y - radius; hi.y = c.y + radius; }
    if (md != 2) { lo.z = c.z - radius; hi.z = c.z + radius; }
    return {min3(lo, hi), max3(lo, hi)};
}

__device__ __forceinline__ uint32_t morton_of(Box frame, f3 p) {
    f3 o = sub3(p, frame.lo);
    if (frame.hi.x > frame.lo.x) o.x /= (frame.hi.x - frame.lo.x);
    if (frame.hi.y > frame.lo.y) o.y /= (frame.hi.y - frame.lo.y);
    if (frame.hi.z > frame.lo.z) o.z /= (frame.hi.z - frame.lo.z);
    o = mul3(1024.0f, o);
    return morton3(f2u_sat(o.x), f2u_sat(o.y), f2u_sat(o.z));
}

// Morton code per primitive + the three 10-bit digit histograms of the sort.
__global__ __launch_bounds__(256) void morton_kernel(const float *__restrict__ tri, uint32_t n,
                                                     const uint32_t *__restrict__ cbounds,
                                                     uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                     uint32_t *__restrict__ hist) {
    __shared__ uint32_t s_hist[kSortPasses * kSortRadix];
    for (int i = threadIdx.x; i < int(kSortPasses * kSortRadix); i += blockDim.x) s_hist[i] = 0u;
    __syncthreads();
    const Box frame = morton_frame(cbounds);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *t = tri + 9ull * i;
        f3 p0{t[0], t[1], t[2]}, p1{t[3], t[4], t[5]}, p2{t[6], t[7], t[8]};
        const uint32_t code = morton_of(frame, box_centroid(tri_box(p0, p1, p2)));
        keys[i] = code;
        vals[i] = i;
        atomicAdd(&s_hist[code & 1023u], 1u);
        atomicAdd(&s_hist[kSortRadix + ((code >> 10) & 1023u)], 1u);
        atomicAdd(&s_hist[2 * kSortRadix + ((code >> 20) & 1023u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < int(kSortPasses * kSortRadix); i += blockDim.x)
        if (s_hist[i]) atomicAdd(&hist[i], s_hist[i]);
}

// Sorted position -> primitive record (box of the triangle, its sorted position
// i: the leaves' primitive order holds sorted positions until leaf_order_kernel
// maps them to input indices), the triangle copied in sorted order, and the
// treelet-start flag (code >> 9 differs from the previous position).
__global__ __launch_bounds__(256) void gather_kernel(const float *__restrict__ tri, uint32_t n,
                                                     const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ vals, Rec *__restrict__ prec,
                                                     float *__restrict__ tri_sorted, uint32_t *__restrict__ flags) {
    // the block's 256 triangles are staged in LDS and written as one contiguous
    // run (16-byte stores for a full block: 9216 bytes, 16-byte aligned)
    __shared__ __align__(16) float s_tri[256 * 9];
    const uint32_t i0 = blockIdx.x * 256u;
    const uint32_t i = i0 + threadIdx.x;
    if (i < n) {
        const uint32_t pid = vals[i] < n ? vals[i] : 0u;  // in range unless the sort failed (reported)
        const float *t = tri + 9ull * pid;
        f3 p0{t[0], t[1], t[2]}, p1{t[3], t[4], t[5]}, p2{t[6], t[7], t[8]};
        prec[i] = make_rec(tri_box(p0, p1, p2), i, 1u);
        float *d = s_tri + 9 * threadIdx.x;
        d[0] = p0.x; d[1] = p0.y; d[2] = p0.z;
        d[3] = p1.x; d[4] = p1.y; d[5] = p1.z;
        d[6] = p2.x; d[7] = p2.y; d[8] = p2.z;
        const uint32_t key = keys[i] >> kTreeletShift;
        flags[i] = (i == 0 || (keys[i - 1] >> kTreeletShift) != key) ? 1u : 0u;
    }
    __syncthreads();
    if (n - i0 >= 256u) {
        float4 *dst = reinterpret_cast<float4 *>(tri_sorted + 9ull * i0);
        const float4 *src = reinterpret_cast<const float4 *>(s_tri);
        for (uint32_t j = threadIdx.x; j < 256u * 9u / 4u; j += 256u) dst[j] = src[j];
    } else {
        const uint32_t cnt = (n - i0) * 9u;
        for (uint32_t j = threadIdx.x; j < cnt; j += 256u) tri_sorted[9ull * i0 + j] = s_tri[j];
    }
}

// Leaf position k: the sorted position m the tree put there becomes the input
// index vals[m], and the triangle is copied from the sorted copy (48-byte
// stride, three 16-byte stores). A treelet's primitives are one contiguous run
// of sorted positions, so these reads stay local where the input-order gather
// would be scattered.
__global__ __launch_bounds__(256) void leaf_order_kernel(const uint32_t *__restrict__ vals,
                                                         const float *__restrict__ tri_sorted, uint32_t n,
                                                         uint32_t *__restrict__ prim_order,
                                                         float *__restrict__ tri_leaf) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t m0 = prim_order[k];
    const uint32_t m = m0 < n ? m0 : 0u;  // m0 < n for every complete tree
    prim_order[k] = vals[m];
    const float *s = tri_sorted + 9ull * m;
    float4 *d = reinterpret_cast<float4 *>(tri_leaf) + 3ull * k;
    d[0] = make_float4(s[0], s[1], s[2], s[3]);
    d[1] = make_float4(s[4], s[5], s[6], s[7]);
    d[2] = make_float4(s[8], 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void treelet_start_kernel(const uint32_t *__restrict__ flags,
                                                            const uint32_t *__restrict__ tidx, uint32_t n,
                                                            uint32_t *__restrict__ tstart) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) tstart[tidx[i]] = i;
}

// Treelet box = union of member boxes (hlbvh.cu:82-99).
__global__ __launch_bounds__(256) void treelet_rec_kernel(const uint32_t *__restrict__ tstart, uint32_t d,
                                                          uint32_t n, const Rec *__restrict__ prec,
                                                          Rec *__restrict__ trec) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d) return;
    const uint32_t s = tstart[t], e = (t + 1 < d) ? tstart[t + 1] : n;
    Box b = box_empty();
    for (uint32_t i = s; i < e; ++i) b = box_union(b, rec_box(prec[i]));
    trec[t] = make_rec(b, s, e - s);
}

__global__ void root_kernel(BuildArgs A, uint32_t num_treelets, uint32_t n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (num_treelets == 1) {
        emit_bottom_root(A, A.trec[0][0], 0);
        return;
    }
    WorkItem it{0, num_treelets, 0, n, 0, {0, 0, 0}};
    if (num_treelets <= kSmallNode) push_item(A.small_top, A.counters + kCntSmallTop, A.cap_small_top, it, A.err);
    else push_item(A.mid_top, A.counters + kCntMidTop, A.cap_mid_top, it, A.err);
}

// ------------------------------------------------- level step for large nodes

// Level kernels read the level's item count (*nitems) and task count (*ntasks)
// from device memory, so that several levels are enqueued without reading them
// back; grids cover host-side upper bounds and the excess exits.

// Start of a level, one thread per accumulator word of the `bound` item slots:
// chunk counts (zero past the level's items, for the scan over the bound), the
// items' accumulators set to their identities, and the next level's list
// counter zeroed (no kernel of this level reads it).
__global__ __launch_bounds__(256) void level_setup_kernel(const WorkItem *items, const uint32_t *nitems,
                                                          uint32_t bound, uint32_t *counts, Acc *acc,
                                                          uint32_t *next_counter) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) *next_counter = 0u;
    const uint32_t i = g / kAccWords, k = g % kAccWords;
    if (i >= bound) return;
    const uint32_t n = *nitems;
    if (k == 0) counts[i] = i < n ? (items[i].end - items[i].begin + kChunk - 1) / kChunk : 0u;
    if (i >= n) return;
    uint32_t v = 0u;
    if (k < 3 || (k >= 6 && k < 9)) v = 0xFFFFFFFFu;  // cb.lo, fb.lo
    if (k >= 12 + kNumBuckets) {
        const uint32_t j = (k - 12 - kNumBuckets) % 6;
        v = j < 3 ? 0xFFFFFFFFu : 0u;
    }
    reinterpret_cast<uint32_t *>(&acc[i])[k] = v;
}

// One thread per task: its item is the last one whose first task is <= t (every
// item has at least one chunk, so task_off is strictly increasing).
__global__ __launch_bounds__(256) void task_fill_kernel(const uint32_t *nitems, uint32_t bound,
                                                        const uint32_t *task_off, const uint32_t *ntasks,
                                                        uint2 *tasks) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *ntasks) return;
    uint32_t lo = 0, hi = min(*nitems, bound);  // task_off holds bound entries  // task_off[lo] <= t < task_off[hi] (task_off[nitems] = ntasks)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (task_off[mid] <= t) lo = mid;
        else hi = mid;
    }
    tasks[t] = make_uint2(lo, t - task_off[lo]);
}

// Block reduce of 6 order-mapped mins and 6 maxes, merged into global atomics.
__device__ __forceinline__ void block_merge_boxes(uint32_t v[12], uint32_t *dst) {
    __shared__ uint32_t s[kChunkBlock / kWave][12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const bool is_min = (k % 6) < 3;
        v[k] = is_min ? wave_min_u32(v[k]) : wave_max_u32(v[k]);
    }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int k = 0; k < 12; ++k) s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool is_min = (k % 6) < 3;
        uint32_t r = s[0][k];
        for (int ww = 1; ww < kChunkBlock / kWave; ++ww) r = is_min ? min(r, s[ww][k]) : max(r, s[ww][k]);
        if (is_min) atomicMin(&dst[k], r);
        else atomicMax(&dst[k], r);
    }
}

__global__ __launch_bounds__(kChunkBlock) void chunk_reduce_kernel(const uint2 *tasks, const uint32_t *ntasks,
                                                                   const WorkItem *items, RecPair recs, Acc *acc) {
    if (blockIdx.x >= *ntasks) return;
    const uint2 task = tasks[blockIdx.x];
    const WorkItem it = items[task.x];
    const Rec *R = pick(recs.buf, it.buf);
    const uint32_t s = it.begin + task.y * kChunk, e = min(it.end, s + kChunk);
    uint32_t v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? 0xFFFFFFFFu : 0u;
    for (uint32_t i = s + threadIdx.x; i < e; i += kChunkBlock) {
        const Box b = rec_box(R[i]);
        const f3 c = box_centroid(b);
        v[0] = min(v[0], f2ord(c.x)); v[1] = min(v[1], f2ord(c.y)); v[2] = min(v[2], f2ord(c.z));
        v[3] = max(v[3], f2ord(c.x)); v[4] = max(v[4], f2ord(c.y)); v[5] = max(v[5], f2ord(c.z));
        v[6] = min(v[6], f2ord(b.lo.x)); v[7] = min(v[7], f2ord(b.lo.y)); v[8] = min(v[8], f2ord(b.lo.z));
        v[9] = max(v[9], f2ord(b.hi.x)); v[10] = max(v[10], f2ord(b.hi.y)); v[11] = max(v[11], f2ord(b.hi.z));
    }
    block_merge_boxes(v, acc[task.x].cb);
}

// SAH bucket statistics per chunk (hlbvh.cu:690-706).
__global__ __launch_bounds__(kChunkBlock) void chunk_buckets_kernel(const uint2 *tasks, const uint32_t *ntasks,
                                                                    const WorkItem *items, RecPair recs, Acc *acc,
                                                                    uint32_t *err) {
    __shared__ uint32_t s_cnt[kNumBuckets];
    __shared__ uint32_t s_box[kNumBuckets][6];
    if (blockIdx.x >= *ntasks) return;
    const uint2 task = tasks[blockIdx.x];
    const WorkItem it = items[task.x];
    const Rec *R = pick(recs.buf, it.buf);
    for (int k = threadIdx.x; k < int(kNumBuckets * 7); k += kChunkBlock) {
        if (k < int(kNumBuckets)) s_cnt[k] = 0u;
        else {
            int j = k - kNumBuckets;
            (&s_box[0][0])[j] = (j % 6) < 3 ? 0xFFFFFFFFu : 0u;
        }
    }
    const Box cb = acc_box(acc[task.x].cb);
    const int axis = box_max_dim(cb);
    const float base = sel3(cb.lo, axis);
    const float span = sel3(cb.hi, axis) - sel3(cb.lo, axis);
    if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) {
        if (threadIdx.x == 0) atomicOr(err, uint32_t(kErrDegenerateCentroids));
        return;
    }
    __syncthreads();
    const uint32_t s = it.begin + task.y * kChunk, e = min(it.end, s + kChunk);
    for (uint32_t i = s + threadIdx.x; i < e; i += kChunkBlock) {
        const Rec r = R[i];
        const Box b = rec_box(r);
        const uint32_t k = bucket_of(sel3(box_centroid(b), axis), base, span);
        atomicAdd(&s_cnt[k], r.b);
        atomicMin(&s_box[k][0], f2ord(b.lo.x)); atomicMin(&s_box[k][1], f2ord(b.lo.y));
        atomicMin(&s_box[k][2], f2ord(b.lo.z));
        atomicMax(&s_box[k][3], f2ord(b.hi.x)); atomicMax(&s_box[k][4], f2ord(b.hi.y));
        atomicMax(&s_box[k][5], f2ord(b.hi.z));
    }
    __syncthreads();
    Acc *a = &acc[task.x];
    for (int k = threadIdx.x; k < int(kNumBuckets * 7); k += kChunkBlock) {
        if (k < int(kNumBuckets)) {
            if (s_cnt[k]) atomicAdd(&a->bcnt[k], s_cnt[k]);
        } else {
            const int j = k - kNumBuckets;
            const uint32_t v = (&s_box[0][0])[j];
            if ((j % 6) < 3) { if (v != 0xFFFFFFFFu) atomicMin(&(&a->bbox[0][0])[j], v); }
            else if (v != 0u) atomicMax(&(&a->bbox[0][0])[j], v);
        }
    }
}

// Split choice of one large node. Top level: one wave per node, lane k holds
// bucket k; inclusive prefix and suffix unions (exact min/max, so equal to the
// reference's per-split loops, hlbvh.cu:711-731) give lane k the cost of
// splitting after bucket k; lane 0 then takes the first strict minimum
// (hlbvh.cu:734-741). Bottom level: the midpoint of the centroid box.
__device__ __forceinline__ float shfl_up_f(float v, int d) { return __shfl_up(v, d, kWave); }
__device__ __forceinline__ float shfl_down_f(float v, int d) { return __shfl_down(v, d, kWave); }

template <bool TOP>
__global__ __launch_bounds__(64) void decide_kernel(const WorkItem *items, const uint32_t *nitems, const Acc *acc,
                                                    Split *splits) {
    __shared__ float s_cost[kNumBuckets];
    const uint32_t item = TOP ? blockIdx.x : blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= *nitems) return;
    const Acc &a = acc[item];
    const Box cb = acc_box(a.cb);
    Split sp{};
    sp.axis = box_max_dim(cb);
    sp.split_val = sel3(box_centroid(cb), sp.axis);
    sp.mode = 1;
    if (!TOP) {
        splits[item] = sp;
        return;
    }
    const int lane = threadIdx.x;
    const bool in = lane < int(kNumBuckets);
    const uint32_t cnt = in ? a.bcnt[lane] : 0u;
    // an untouched bucket holds the atomic identities, which decode to NaN: use the empty box
    Box b = cnt > 0 ? acc_box(a.bbox[lane]) : box_empty();
    // inclusive prefix (left) and suffix (right) unions and counts over buckets
    float pl[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
    float sf[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
    uint32_t pc = cnt, sc = cnt;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const bool up = lane >= d;
        const bool dn = lane + d < int(kNumBuckets);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float u = shfl_up_f(pl[k], d);
            const float w = shfl_down_f(sf[k], d);
            if (up) pl[k] = k < 3 ? smin(u, pl[k]) : smax(u, pl[k]);
            if (dn) sf[k] = k < 3 ? smin(sf[k], w) : smax(sf[k], w);
        }
        const uint32_t uc = __shfl_up(pc, d, kWave), wc = __shfl_down(sc, d, kWave);
        if (up) pc += uc;
        if (dn) sc += wc;
    }
    // right side of split k = suffix starting at bucket k + 1
    float r6[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) r6[k] = shfl_down_f(sf[k], 1);
    const uint32_t rc = __shfl_down(sc, 1, kWave);
    const float total_sa = box_area(acc_box(a.fb));
    if (lane < int(kNumBuckets) - 1) {
        const Box bl{{pl[0], pl[1], pl[2]}, {pl[3], pl[4], pl[5]}};
        const Box br{{r6[0], r6[1], r6[2]}, {r6[3], r6[4], r6[5]}};
        s_cost[lane] = sah_cost((float(pc) * box_area(bl) + float(rc) * box_area(br)) / total_sa);
    }
    __syncthreads();
    uint32_t split = 0;
    if (lane == 0) {
        float best = s_cost[0];
        for (int k = 1; k < int(kNumBuckets) - 1; ++k)
            if (s_cost[k] < best) { best = s_cost[k]; split = k; }
    }
    split = __shfl(split, 0, kWave);
    const uint32_t lp = __shfl(pc, int(split), kWave);
    const uint32_t rp = __shfl(rc, int(split), kWave);
    if (lane == 0) {
        sp.bucket = split;
        sp.base = sel3(cb.lo, sp.axis);
        sp.span = sel3(cb.hi, sp.axis) - sel3(cb.lo, sp.axis);
        sp.left_prims = lp;
        sp.mode = (lp == 0 || rp == 0) ? 1u : 0u;  // midpoint fallback (hlbvh.cu:758-777)
        splits[item] = sp;
    }
}

template <bool TOP>
__device__ __forceinline__ bool goes_left(const Split &sp, const Box &b) {
    const float c = sel3(box_centroid(b), sp.axis);
    if (TOP && sp.mode == 0) return bucket_of(c, sp.base, sp.span) <= sp.bucket;
    return c < sp.split_val;
}

template <bool TOP>
__global__ __launch_bounds__(kChunkBlock) void chunk_count_kernel(const uint2 *tasks, const uint32_t *ntasks,
                                                                  const WorkItem *items, RecPair recs,
                                                                  const Split *splits, uint32_t *task_left,
                                                                  uint32_t *task_lprims) {
    __shared__ uint32_t s_w[kChunkBlock / kWave];
    __shared__ uint32_t s_p[kChunkBlock / kWave];
    if (blockIdx.x >= *ntasks) {  // past the level's tasks: zeros for the scans over the bound
        if (threadIdx.x == 0) {
            task_left[blockIdx.x] = 0u;
            task_lprims[blockIdx.x] = 0u;
        }
        return;
    }
    const uint2 task = tasks[blockIdx.x];
    const WorkItem it = items[task.x];
    const Split sp = splits[task.x];
    const Rec *R = pick(recs.buf, it.buf);
    const uint32_t s = it.begin + task.y * kChunk, e = min(it.end, s + kChunk);
    uint32_t cnt = 0, prims = 0;
    for (uint32_t i = s + threadIdx.x; i < e; i += kChunkBlock) {
        const Rec r = R[i];
        if (goes_left<TOP>(sp, rec_box(r))) { ++cnt; prims += r.b; }
    }
    cnt = wave_sum(cnt);
    prims = wave_sum(prims);
    if (__lane_id() == 0) { s_w[threadIdx.x / kWave] = cnt; s_p[threadIdx.x / kWave] = prims; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0, p = 0;
        for (int w = 0; w < kChunkBlock / kWave; ++w) { t += s_w[w]; p += s_p[w]; }
        task_left[blockIdx.x] = t;
        task_lprims[blockIdx.x] = p;
    }
}

// Stable partition of one chunk into the other buffer.
template <bool TOP>
__global__ __launch_bounds__(kChunkBlock) void chunk_scatter_kernel(const uint2 *tasks, const uint32_t *ntasks,
                                                                    const WorkItem *items, RecPair recs,
                                                                    const Split *splits, const uint32_t *task_off,
                                                                    const uint32_t *left_off) {
    __shared__ uint32_t s_waves[kChunkBlock / kWave];
    const uint32_t t = blockIdx.x;
    if (t >= *ntasks) return;
    const uint2 task = tasks[t];
    const WorkItem it = items[task.x];
    const Split sp = splits[task.x];
    const Rec *R = pick(recs.buf, it.buf);
    Rec *W = pick(recs.buf, 1 - it.buf);
    const uint32_t first_task = task_off[task.x];
    const uint32_t nchunks = (it.end - it.begin + kChunk - 1) / kChunk;
    const uint32_t left_total = left_off[first_task + nchunks] - left_off[first_task];
    const uint32_t left_before = left_off[t] - left_off[first_task];
    const uint32_t s = it.begin + task.y * kChunk;
    const uint32_t right_before = (s - it.begin) - left_before;
    const uint32_t e = min(it.end, s + kChunk);
    Rec r[kChunkItems];
    uint32_t mask = 0, nleft = 0;
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) {
        const uint32_t i = s + threadIdx.x * kChunkItems + k;
        if (i < e) {
            r[k] = R[i];
            if (goes_left<TOP>(sp, rec_box(r[k]))) { mask |= 1u << k; ++nleft; }
        }
    }
    const uint32_t nvalid = (s + threadIdx.x * kChunkItems >= e)
                                ? 0u
                                : min(uint32_t(kChunkItems), e - (s + threadIdx.x * kChunkItems));
    uint32_t tot;
    const uint32_t lpre = block_exclusive_add<kChunkBlock>(nleft, s_waves, &tot);
    const uint32_t rpre = threadIdx.x * kChunkItems - lpre;  // elements before this thread minus lefts
    uint32_t li = 0, ri = 0;
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) {
        if (uint32_t(k) >= nvalid) break;
        if (mask & (1u << k)) W[it.begin + left_before + lpre + li++] = r[k];
        else W[it.begin + left_total + right_before + rpre + ri++] = r[k];
    }
}

template <bool TOP>
__global__ __launch_bounds__(64) void children_kernel(BuildArgs A, const WorkItem *items, const uint32_t *nitems,
                                                      const Split *splits, const Acc *acc,
                                                      const uint32_t *task_off, const uint32_t *left_off,
                                                      const uint32_t *lprims_off, WorkItem *next_large,
                                                      uint32_t *next_counter, uint32_t *levels) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *nitems;
    if (i == 0 && n > 0) atomicAdd(levels, 1u);  // a level that had items
    if (i >= n) return;
    const WorkItem it = items[i];
    const Split sp = splits[i];
    const uint32_t size = it.end - it.begin;
    const uint32_t first_task = task_off[i];
    const uint32_t nchunks = (size + kChunk - 1) / kChunk;
    const uint32_t L = left_off[first_task + nchunks] - left_off[first_task];
    const Box fb = acc_box(acc[i].fb);
    Rec *const(&recs)[2] = TOP ? A.trec : A.prec;
    if (L == 0 || L == size) {
        if (TOP) {  // unreachable for finite input: both SAH end buckets are occupied
            atomicOr(A.err, uint32_t(kErrSplit));
            return;
        }
        write_leaf(A.nodes, it.node, fb, it.begin, size);  // all centroids on one side
        const Rec *R = pick(recs, it.buf);
        for (uint32_t k = it.begin; k < it.end; ++k) A.prim_order[k] = R[k].a;
        return;
    }
    const uint32_t lp = TOP ? (lprims_off[first_task + nchunks] - lprims_off[first_task]) : L;
    write_interior(A.nodes, it.node, fb, lp, sp.axis);
    const uint32_t nb = 1 - it.buf;
    WorkItem ch[2] = {{it.begin, it.begin + L, it.node + 1, lp, nb, {0, 0, 0}},
                      {it.begin + L, it.end, it.node + 2 * lp, it.prims - lp, nb, {0, 0, 0}}};
    for (int c = 0; c < 2; ++c) {
        WorkItem cw = ch[c];
        const uint32_t cs = cw.end - cw.begin;
        // 0 = done here, 1 = small list, 2 = next large list, 3 = small bottom, 4 = large bottom
        int cat = 0;
        if (cs == 1) {
            const Rec r = pick(recs, nb)[cw.begin];
            if (TOP) {  // a single treelet: root of a bottom-level build over its primitives
                cw = WorkItem{r.a, r.a + r.b, cw.node, r.b, 0, {0, 0, 0}};
                if (r.b == 1) {
                    const Rec p = A.prec[0][r.a];
                    write_leaf(A.nodes, cw.node, rec_box(p), r.a, 1);
                    A.prim_order[r.a] = p.a;
                } else {
                    cat = r.b <= kSmallNode ? 3 : 4;
                }
            } else {
                write_leaf(A.nodes, cw.node, rec_box(r), cw.begin, 1);
                A.prim_order[cw.begin] = r.a;
            }
        } else {
            cat = cs <= kSmallNode ? 1 : 2;
        }
        wave_push(cat == 1, cw, TOP ? A.small_top : A.small_bot,
                  A.counters + (TOP ? kCntSmallTop : kCntSmallBot), TOP ? A.cap_small_top : A.cap_small_bot, A.err);
        wave_push(cat == 2, cw, next_large, next_counter, TOP ? A.cap_large_top : A.cap_large_bot, A.err);
        wave_push(cat == 3, cw, A.small_bot, A.counters + kCntSmallBot, A.cap_small_bot, A.err);
        wave_push(cat == 4, cw, A.large_bot[0], A.counters + kCntLargeBot0, A.cap_large_bot, A.err);
    }
}



// ------------------------------------------------ SAH price of the 23 splits
//
// Lane k (< 24) holds bucket k: its primitive count, treelet count and box (the
// empty box when unused); lane k also holds bucket 23 - k in the `r` arguments.
// Inclusive prefixes over lanes 0..23 by DPP (row shifts 1, 2, 4, 8, then row 0's
// last lane broadcast into row 1): the forward scan gives the left side of split k
// at lane k, the scan of the reversed buckets the right side (buckets k + 1..23)
// at lane 22 - k. Cost of split k as hlbvh.cu:711-731, the first strict minimum as
// :733-741; the min / max / sum of the unions are exact, so the order in which the
// scans combine them changes no value.
struct SplitPrice {
    uint32_t split, left_prims, right_prims, left_treelets;
};

template <typename Op>
__device__ __forceinline__ uint32_t dpp_prefix24(uint32_t v, uint32_t ident, Op op) {
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(ident), int(v), 0x111, 0xF, 0xF, false)));  // row_shr:1
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(ident), int(v), 0x112, 0xF, 0xF, false)));  // row_shr:2
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(ident), int(v), 0x114, 0xF, 0xF, false)));  // row_shr:4
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(ident), int(v), 0x118, 0xF, 0xF, false)));  // row_shr:8
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(int(ident), int(v), 0x142, 0xA, 0xF, false)));  // row_bcast:15
    return v;
}

__device__ __forceinline__ void dpp_prefix_box(uint32_t w[6]) {  // order-mapped lo (min) and hi (max)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[k] = dpp_prefix24(w[k], 0xFFFFFFFFu, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
        w[3 + k] = dpp_prefix24(w[3 + k], 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    }
}

// Every lane of the wave calls it. Boxes are order-mapped (f2ord), identities for
// empty buckets.
__device__ __forceinline__ SplitPrice price_splits(uint32_t cnt, uint32_t tcnt, const uint32_t box[6],
                                                   uint32_t rcnt, const uint32_t rbox[6], float total_sa) {
    const uint32_t lane = __lane_id();
    auto add = [](uint32_t a, uint32_t b) { return a + b; };
    uint32_t pl[6], sr[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        pl[k] = box[k];
        sr[k] = rbox[k];
    }
    dpp_prefix_box(pl);
    dpp_prefix_box(sr);
    const uint32_t pc = dpp_prefix24(cnt, 0u, add), tc = dpp_prefix24(tcnt, 0u, add);
    const uint32_t sc = dpp_prefix24(rcnt, 0u, add);
    const int src = int(kNumBuckets) - 2 - int(lane);  // right side of split `lane` (lanes 0..22)
    uint32_t r6[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) r6[k] = __shfl(sr[k], src < 0 ? 0 : src, kWave);
    const uint32_t rc = __shfl(sc, src < 0 ? 0 : src, kWave);
    uint32_t key = 0xFFFFFFFFu;
    if (lane < kNumBuckets - 1) {
        const Box bl = acc_box(pl), br = acc_box(r6);
        const float cost = sah_cost((float(pc) * box_area(bl) + float(rc) * box_area(br)) / total_sa);
        key = f2ord(cost);
    }
    const uint32_t kmin = wave_min_u32_all(key);
    SplitPrice sp;
    sp.split = uint32_t(__builtin_ctzll(__ballot(lane < kNumBuckets - 1 && key == kmin)));
    sp.left_prims = __shfl(pc, int(sp.split), kWave);
    sp.right_prims = __shfl(rc, int(sp.split), kWave);
    sp.left_treelets = __shfl(tc, int(sp.split), kWave);
    return sp;
}

// ------------------------------------------- fused top level (> 64 treelets)
//
// Four launches per level instead of the generic level path's eleven, and two
// passes over the treelet records instead of four:
//  1. top_bucket_kernel, one workgroup per 2048-treelet chunk of a node: each
//     treelet's SAH bucket (hlbvh.cu:690-706), kept as one byte for the scatter,
//     the chunk's treelet count per bucket, and the bucket prim counts and boxes
//     merged into the node's accumulator.
//  2. top_decide_kernel, one wave per node: prices the 23 splits (prefix and
//     suffix unions over the buckets, first strict minimum: hlbvh.cu:711-741) and
//     turns the chunks' counts into each chunk's offset among the left side.
//  3. top_scatter_kernel: the stable partition of each chunk into the other
//     buffer by its bucket bytes, merging the two children's centroid and full
//     boxes into the node's child boxes on the way.
//  4. top_children_kernel, one thread per node: the interior node, then each child
//     becomes a bottom-level root (one treelet), a node for the wave finisher
//     (<= 64 treelets) or a node of the next level, whose boxes come from the
//     child boxes and whose chunk tasks it appends.
constexpr uint32_t kTopInvalid = 0xFFFFFFFFu;

struct TopItem {
    uint32_t begin, end, node, prims;
    uint32_t buf, first_task, ntasks, pad;
};

struct TopAcc {
    uint32_t cb[6], fb[6];        // the node's centroid and full box (order-mapped), from its parent
    uint32_t bcnt[kNumBuckets];   // prims per bucket
    uint32_t bbox[kNumBuckets][6];
    uint32_t child[2][12];        // the children's centroid and full boxes, merged by the scatter
    uint32_t done;                // chunks finished with their buckets
    uint32_t split;               // last bucket on the left side, kTopInvalid until priced
    uint32_t lp, lcount, axis;    // left prims, left treelets, split axis
    uint32_t pad[3];
};

static_assert(offsetof(TopAcc, bcnt) == 12 * 4 && offsetof(TopAcc, bbox) == (12 + kNumBuckets) * 4 &&
                  offsetof(TopAcc, child) == (12 + 7 * kNumBuckets) * 4 &&
                  offsetof(TopAcc, done) == (12 + 7 * kNumBuckets + 24) * 4 &&
                  offsetof(TopAcc, split) == (12 + 7 * kNumBuckets + 25) * 4,
              "TopAcc word layout");

struct TopLevel {
    TopItem *items[2];
    TopAcc *acc[2];
    uint2 *tasks[2];
    uint32_t *hist;   // per task: treelets per bucket
    uint32_t *left;   // per task: treelets of the node's left side in earlier chunks
    uint8_t *bkt;     // per treelet position: its bucket at the current level
    uint32_t cap_items, cap_tasks, num;  // list capacities; treelet count
};

// A level's task and node counts, held to the list capacities (an overflow is
// reported by the kernel that appended; the build fails, reading nothing past the lists).
__device__ __forceinline__ uint32_t top_count(const uint32_t *counters, int slot, uint32_t cap) {
    return min(counters[slot], cap);
}

__device__ __forceinline__ bool top_item_ok(const TopLevel &TL, const TopItem &it) {
    return it.begin < it.end && it.end <= TL.num && it.buf < 2u;
}

__device__ __forceinline__ void top_acc_init(TopAcc &a, const uint32_t box12[12]) {
    for (int k = 0; k < 12; ++k) (k < 6 ? a.cb[k] : a.fb[k - 6]) = box12[k];
    for (int k = 0; k < int(kNumBuckets); ++k) {
        a.bcnt[k] = 0u;
        for (int j = 0; j < 6; ++j) a.bbox[k][j] = j < 3 ? 0xFFFFFFFFu : 0u;
    }
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < 12; ++k) a.child[c][k] = (k % 6) < 3 ? 0xFFFFFFFFu : 0u;
    a.done = 0u;
    a.split = kTopInvalid;
    a.lp = a.lcount = a.axis = 0u;
}

// The root node of the top level: its item, chunk tasks and accumulator (boxes
// identities, merged next by top_root_box_kernel).
__global__ void top_root_kernel(TopLevel TL, uint32_t num_treelets, uint32_t n, uint32_t *counters) {
    if (blockIdx.x != 0) return;
    const uint32_t ntasks = (num_treelets + kChunk - 1) / kChunk;
    for (uint32_t c = threadIdx.x; c < ntasks; c += blockDim.x) TL.tasks[0][c] = make_uint2(0u, c);
    if (threadIdx.x == 0) {
        TL.items[0][0] = TopItem{0u, num_treelets, 0u, n, 0u, 0u, ntasks, 0u};
        uint32_t id[12];
        for (int k = 0; k < 12; ++k) id[k] = (k % 6) < 3 ? 0xFFFFFFFFu : 0u;
        top_acc_init(TL.acc[0][0], id);
        counters[kCntTopItems0] = 1u;
        counters[kCntTopTasks0] = ntasks;
        counters[kCntTopItems1] = 0u;
        counters[kCntTopTasks1] = 0u;
    }
}

// The root's centroid and full box over all treelet records.
__global__ __launch_bounds__(kChunkBlock) void top_root_box_kernel(const Rec *__restrict__ R, uint32_t m,
                                                                   TopLevel TL) {
    uint32_t v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? 0xFFFFFFFFu : 0u;
    for (uint32_t i = blockIdx.x * kChunkBlock + threadIdx.x; i < m; i += gridDim.x * kChunkBlock) {
        const Box b = rec_box(R[i]);
        const f3 c = box_centroid(b);
        v[0] = min(v[0], f2ord(c.x)); v[1] = min(v[1], f2ord(c.y)); v[2] = min(v[2], f2ord(c.z));
        v[3] = max(v[3], f2ord(c.x)); v[4] = max(v[4], f2ord(c.y)); v[5] = max(v[5], f2ord(c.z));
        v[6] = min(v[6], f2ord(b.lo.x)); v[7] = min(v[7], f2ord(b.lo.y)); v[8] = min(v[8], f2ord(b.lo.z));
        v[9] = max(v[9], f2ord(b.hi.x)); v[10] = max(v[10], f2ord(b.hi.y)); v[11] = max(v[11], f2ord(b.hi.z));
    }
    block_merge_boxes(v, TL.acc[0][0].cb);  // cb[6] and fb[6] are contiguous
}

__global__ __launch_bounds__(kChunkBlock) void top_bucket_kernel(TopLevel TL, uint32_t p, RecPair recs,
                                                                 uint32_t *counters, uint32_t *err) {
    __shared__ uint32_t s_cnt[kNumBuckets], s_pr[kNumBuckets], s_box[kNumBuckets][6];
    const uint32_t q = 1u - p;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the next level's lists start empty
        counters[kCntTopItems0 + q] = 0u;
        counters[kCntTopTasks0 + q] = 0u;
    }
    const uint32_t t = blockIdx.x;
    if (t >= top_count(counters, kCntTopTasks0 + p, TL.cap_tasks)) return;
    const uint2 task = TL.tasks[p][t];
    if (task.x >= TL.cap_items) return;
    const TopItem it = TL.items[p][task.x];
    if (!top_item_ok(TL, it)) return;
    TopAcc &A = TL.acc[p][task.x];
    const Rec *R = pick(recs.buf, it.buf);
    const Box cb = acc_box(A.cb);
    const int axis = box_max_dim(cb);
    const float base = sel3(cb.lo, axis);
    const float span = sel3(cb.hi, axis) - sel3(cb.lo, axis);
    if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) {  // the node stays unpriced: no children
        if (threadIdx.x == 0) atomicOr(err, uint32_t(kErrDegenerateCentroids));
        return;
    }
    for (int k = threadIdx.x; k < int(kNumBuckets * 8); k += kChunkBlock) {
        if (k < int(kNumBuckets)) s_cnt[k] = 0u;
        else if (k < int(2 * kNumBuckets)) s_pr[k - kNumBuckets] = 0u;
        else {
            const int j = k - 2 * kNumBuckets;
            (&s_box[0][0])[j] = (j % 6) < 3 ? 0xFFFFFFFFu : 0u;
        }
    }
    __syncthreads();
    // Treelets next to each other in Morton order mostly share a bucket: each wave
    // reduces the lanes of its first two buckets across the wave and one lane
    // merges each result into LDS (instead of many same-address atomics); lanes
    // of further buckets (deep levels, where a wave spans many) merge their own.
    const uint32_t s = it.begin + task.y * kChunk, e = min(it.end, s + kChunk);
    for (uint32_t i0 = s; i0 < e; i0 += kChunkBlock) {  // the same trip count in every thread
        const uint32_t i = i0 + threadIdx.x;
        const bool valid = i < e;
        Rec r{};
        uint32_t k = 0;
        if (valid) {
            r = R[i];
            k = bucket_of(sel3(box_centroid(rec_box(r)), axis), base, span);
            TL.bkt[i] = uint8_t(k);
        }
        uint64_t todo = __ballot(valid);
        for (int round = 0; round < 2 && todo; ++round) {
            const int lead = __builtin_ctzll(todo);
            const uint32_t kb = __shfl(k, lead, kWave);
            const bool in = valid && k == kb;
            const uint64_t peers = __ballot(in);
            todo &= ~peers;
            const uint32_t pr = wave_sum_all(in ? r.b : 0u);
            uint32_t bx[6];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                bx[d] = wave_min_u32_all(in ? f2ord(r.lo[d]) : 0xFFFFFFFFu);
                bx[3 + d] = wave_max_u32_all(in ? f2ord(r.hi[d]) : 0u);
            }
            if (int(__lane_id()) == lead) {
                atomicAdd(&s_cnt[kb], uint32_t(__popcll(peers)));
                atomicAdd(&s_pr[kb], pr);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    atomicMin(&s_box[kb][d], bx[d]);
                    atomicMax(&s_box[kb][3 + d], bx[3 + d]);
                }
            }
        }
        if ((todo >> __lane_id()) & 1u) {
            atomicAdd(&s_cnt[k], 1u);
            atomicAdd(&s_pr[k], r.b);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                atomicMin(&s_box[k][d], f2ord(r.lo[d]));
                atomicMax(&s_box[k][3 + d], f2ord(r.hi[d]));
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < int(kNumBuckets * 7); k += kChunkBlock) {
        if (k < int(kNumBuckets)) {
            TL.hist[size_t(t) * kNumBuckets + k] = s_cnt[k];
            if (s_pr[k]) atomicAdd(&A.bcnt[k], s_pr[k]);
        } else {
            const int j = k - kNumBuckets;
            if (s_cnt[j / 6] == 0u) continue;
            const uint32_t v = (&s_box[0][0])[j];
            if ((j % 6) < 3) atomicMin(&(&A.bbox[0][0])[j], v);
            else atomicMax(&(&A.bbox[0][0])[j], v);
        }
    }
}

// One wave per node: prices the 23 splits from the node's bucket counts and boxes
// (lane k = bucket k, as decide_kernel), then turns its chunks' per-bucket counts
// into each chunk's offset among the left-side treelets.
__global__ __launch_bounds__(64) void top_decide_kernel(TopLevel TL, uint32_t p, const uint32_t *counters,
                                                        uint32_t *err) {
    const uint32_t item = blockIdx.x;
    if (item >= top_count(counters, kCntTopItems0 + p, TL.cap_items)) return;
    const TopItem it = TL.items[p][item];
    if (!top_item_ok(TL, it) || uint64_t(it.first_task) + it.ntasks > TL.cap_tasks) return;
    TopAcc &A = TL.acc[p][item];
    const Box cb = acc_box(A.cb);
    const int axis = box_max_dim(cb);
    if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) return;  // reported by top_bucket_kernel; no children
    const int lane = threadIdx.x;
    const bool in = lane < int(kNumBuckets);
    const int fl = in ? lane : 0, rl = in ? int(kNumBuckets) - 1 - lane : 0;  // the bucket, reversed
    const uint32_t cnt = in ? A.bcnt[fl] : 0u, rcnt = in ? A.bcnt[rl] : 0u;
    uint32_t box[6], rbox[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        // an untouched bucket holds the atomic identities
        box[k] = cnt > 0 ? A.bbox[fl][k] : (k < 3 ? 0xFFFFFFFFu : 0u);
        rbox[k] = rcnt > 0 ? A.bbox[rl][k] : (k < 3 ? 0xFFFFFFFFu : 0u);
    }
    const SplitPrice price = price_splits(cnt, 0u, box, rcnt, rbox, box_area(acc_box(A.fb)));
    const uint32_t split = price.split, lp = price.left_prims, rp = price.right_prims;
    // both end buckets hold a treelet (the least and the greatest centroid), so
    // neither side is empty; the midpoint fallback (hlbvh.cu:758-777) is unreachable
    if (lp == 0u || rp == 0u) {
        if (lane == 0) atomicOr(err, uint32_t(kErrSplit));
        return;
    }
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < it.ntasks; c0 += kWave) {
        const uint32_t c = c0 + lane;
        uint32_t l = 0;
        if (c < it.ntasks) {  // the chunk's 24 counts in six independent 16-byte loads
            const uint4 *h = reinterpret_cast<const uint4 *>(TL.hist + size_t(it.first_task + c) * kNumBuckets);
            uint4 q[kNumBuckets / 4];
#pragma unroll
            for (int j = 0; j < int(kNumBuckets / 4); ++j) q[j] = h[j];
#pragma unroll
            for (int j = 0; j < int(kNumBuckets / 4); ++j) {
                l += (4 * j + 0 <= int(split)) ? q[j].x : 0u;
                l += (4 * j + 1 <= int(split)) ? q[j].y : 0u;
                l += (4 * j + 2 <= int(split)) ? q[j].z : 0u;
                l += (4 * j + 3 <= int(split)) ? q[j].w : 0u;
            }
        }
        const uint32_t inc = wave_inclusive_add(l);
        if (c < it.ntasks) TL.left[it.first_task + c] = carry + inc - l;
        carry += __shfl(inc, kWave - 1, kWave);
    }
    if (lane == 0) {
        A.lcount = carry;
        A.lp = lp;
        A.axis = uint32_t(axis);
        A.split = split;
    }
}

// Striped over the chunk (item k*256 + thread), so that every load and store
// instruction covers consecutive records: a treelet's rank among the chunk's
// left side is the left count of the rows before its own, plus that of the
// earlier waves and lanes of its row.
__global__ __launch_bounds__(kChunkBlock) void top_scatter_kernel(TopLevel TL, uint32_t p, RecPair recs,
                                                                  const uint32_t *counters) {
    constexpr int kWaves = kChunkBlock / kWave;
    __shared__ uint32_t s_rows[kChunkItems][kWaves];  // left treelets per row and wave
    const uint32_t t = blockIdx.x;
    if (t >= top_count(counters, kCntTopTasks0 + p, TL.cap_tasks)) return;
    const uint2 task = TL.tasks[p][t];
    if (task.x >= TL.cap_items) return;
    const TopItem it = TL.items[p][task.x];
    if (!top_item_ok(TL, it)) return;
    TopAcc &A = TL.acc[p][task.x];
    const uint32_t split = A.split;
    if (split == kTopInvalid) return;
    const Rec *R = pick(recs.buf, it.buf);
    Rec *W = pick(recs.buf, 1 - it.buf);
    const uint32_t s = it.begin + task.y * kChunk, e = min(it.end, s + kChunk);
    const uint32_t left_total = A.lcount;
    const uint32_t left_before = TL.left[t];
    const uint32_t right_before = (s - it.begin) - left_before;
    const int w = threadIdx.x / kWave;
    Rec r[kChunkItems];
    uint32_t lbits = 0;
    uint32_t v[2][12];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < 12; ++k) v[c][k] = ((k % 6) < 3) ? 0xFFFFFFFFu : 0u;
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) {
        const uint32_t i = s + k * kChunkBlock + threadIdx.x;
        bool left = false;
        if (i < e) {
            r[k] = R[i];
            left = TL.bkt[i] <= split;
            const Box b = rec_box(r[k]);
            const f3 c = box_centroid(b);
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
                if ((sd == 0) != left) continue;
                uint32_t *q = v[sd];
                q[0] = min(q[0], f2ord(c.x)); q[1] = min(q[1], f2ord(c.y)); q[2] = min(q[2], f2ord(c.z));
                q[3] = max(q[3], f2ord(c.x)); q[4] = max(q[4], f2ord(c.y)); q[5] = max(q[5], f2ord(c.z));
                q[6] = min(q[6], f2ord(b.lo.x)); q[7] = min(q[7], f2ord(b.lo.y)); q[8] = min(q[8], f2ord(b.lo.z));
                q[9] = max(q[9], f2ord(b.hi.x)); q[10] = max(q[10], f2ord(b.hi.y)); q[11] = max(q[11], f2ord(b.hi.z));
            }
        }
        lbits |= left ? (1u << k) : 0u;
        const uint64_t m = __ballot(left);
        if (__lane_id() == 0) s_rows[k][w] = __popcll(m);
    }
    __syncthreads();
    uint32_t row_base = 0;  // left treelets in the rows before k
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) {
        uint32_t row_total = 0, before = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) {
            const uint32_t c = s_rows[k][ww];
            row_total += c;
            before += ww < w ? c : 0u;
        }
        const bool left = (lbits >> k) & 1u;
        const uint64_t m = __ballot(left);
        const uint32_t pos = k * kChunkBlock + threadIdx.x;  // position in the chunk
        const uint32_t lefts_before = row_base + before + __popcll(m & lanemask_lt());
        if (s + pos < e) {
            if (left) W[it.begin + left_before + lefts_before] = r[k];
            else W[it.begin + left_total + right_before + (pos - lefts_before)] = r[k];
        }
        row_base += row_total;
    }
    block_merge_boxes(v[0], A.child[0]);
    __syncthreads();  // block_merge_boxes reuses its LDS staging
    block_merge_boxes(v[1], A.child[1]);
}

// One thread per node: the interior node, then its children. Next-level nodes are
// appended with one atomic per wave (for their slots and for their chunk tasks);
// each gets its boxes, no split yet and its tasks here, its bucket and child
// accumulators from top_acc_reset_kernel.
__global__ __launch_bounds__(64) void top_children_kernel(BuildArgs A, TopLevel TL, uint32_t p, uint32_t *levels) {
    const uint32_t q = 1u - p;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t *counters = A.counters;
    const uint32_t n = top_count(counters, kCntTopItems0 + p, TL.cap_items);
    if (i == 0 && n > 0) atomicAdd(levels, 1u);
    bool ok = i < n;
    TopItem it{};
    uint32_t L = 0, lp = 0;
    if (ok) {
        it = TL.items[p][i];
        const TopAcc &Ac = TL.acc[p][i];
        ok = Ac.split != kTopInvalid && top_item_ok(TL, it);
        if (ok) {
            L = Ac.lcount;
            lp = Ac.lp;
            if (L == 0 || L == it.end - it.begin) {  // unreachable (both end buckets are occupied)
                atomicOr(A.err, uint32_t(kErrSplit));
                ok = false;
            } else {
                write_interior(A.nodes, it.node, acc_box(Ac.fb), lp, Ac.axis);
            }
        }
    }
    const uint32_t nb = 1 - it.buf;
    for (int c = 0; c < 2; ++c) {  // every lane of the wave takes part (the appends ballot)
        WorkItem cw{};
        if (ok) cw = c == 0 ? WorkItem{it.begin, it.begin + L, it.node + 1, lp, nb, {0, 0, 0}}
                            : WorkItem{it.begin + L, it.end, it.node + 2 * lp, it.prims - lp, nb, {0, 0, 0}};
        const uint32_t cs = cw.end - cw.begin;
        const bool large = ok && cs > kMidNode;
        if (ok && cs == 1) emit_bottom_root(A, A.trec[nb][cw.begin], cw.node);
        wave_push(ok && cs > 1 && cs <= kSmallNode, cw, A.small_top, counters + kCntSmallTop, A.cap_small_top, A.err);
        wave_push(ok && cs > kSmallNode && cs <= kMidNode, cw, A.mid_top, counters + kCntMidTop, A.cap_mid_top, A.err);
        const uint32_t nt = large ? (cs + kChunk - 1) / kChunk : 0u;
        const uint32_t inc_i = wave_inclusive_add(large ? 1u : 0u), inc_t = wave_inclusive_add(nt);
        uint32_t base_i = 0, base_t = 0;
        if (__lane_id() == kWave - 1 && inc_i > 0) {
            base_i = atomicAdd(counters + kCntTopItems0 + q, inc_i);
            base_t = atomicAdd(counters + kCntTopTasks0 + q, inc_t);
        }
        base_i = __shfl(base_i, kWave - 1, kWave);
        base_t = __shfl(base_t, kWave - 1, kWave);
        const uint32_t slot = base_i + inc_i - 1, first = base_t + inc_t - nt;
        const bool fits = large && slot < TL.cap_items && uint64_t(first) + nt <= TL.cap_tasks;
        if (large && !fits) atomicOr(A.err, uint32_t(kErrCapacity));
        if (fits) {
            TL.items[q][slot] = TopItem{cw.begin, cw.end, cw.node, cw.prims, nb, first, nt, 0u};
            TopAcc &Nx = TL.acc[q][slot];
            const TopAcc &Ac = TL.acc[p][i];
            for (int k = 0; k < 6; ++k) {
                Nx.cb[k] = Ac.child[c][k];
                Nx.fb[k] = Ac.child[c][6 + k];
            }
            Nx.split = kTopInvalid;
        }
        // each node's chunk tasks: a few by its own lane, many (near the root: hundreds)
        // by the whole wave
        if (fits && nt <= 8)
            for (uint32_t k = 0; k < nt; ++k) TL.tasks[q][first + k] = make_uint2(slot, k);
        uint64_t todo = __ballot(fits && nt > 8);
        while (todo) {
            const int l = __builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t f = __shfl(first, l, kWave), cnt = __shfl(nt, l, kWave), sl = __shfl(slot, l, kWave);
            for (uint32_t k = __lane_id(); k < cnt; k += kWave) TL.tasks[q][f + k] = make_uint2(sl, k);
        }
    }
}

// The bucket and child accumulators of the next level's nodes: identities.
__global__ __launch_bounds__(256) void top_acc_reset_kernel(TopLevel TL, uint32_t q, const uint32_t *counters) {
    constexpr uint32_t kFirst = 12, kLast = 12 + 7 * kNumBuckets + 24;  // bcnt .. child
    constexpr uint32_t kSpan = kLast - kFirst;
    const uint32_t n = top_count(counters, kCntTopItems0 + q, TL.cap_items);
    const uint64_t total = uint64_t(n) * kSpan;
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
         g += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t item = uint32_t(g / kSpan), k = kFirst + uint32_t(g % kSpan);
        uint32_t v = 0u;
        if (k >= 12 + kNumBuckets && k < 12 + 7 * kNumBuckets) v = ((k - 12 - kNumBuckets) % 6) < 3 ? 0xFFFFFFFFu : 0u;
        else if (k >= 12 + 7 * kNumBuckets) v = ((k - 12 - 7 * kNumBuckets) % 6) < 3 ? 0xFFFFFFFFu : 0u;
        reinterpret_cast<uint32_t *>(&TL.acc[q][item])[k] = v;
    }
}

// ---------------------------------------- wave finisher for 65..2048 treelets
//
// A top-level node of at most kMidNode treelets is split, together with its
// whole subtree down to nodes of <= 64 treelets, by one wave, one node after
// another from a stack in LDS: per split a bounds pass, a bucket pass
// (hlbvh.cu:690-706), the price of the 23 splits (first strict minimum,
// hlbvh.cu:711-741) and a stable partition into the other record buffer, 64
// positions a row. The waves of a workgroup work independently (no workgroup
// barriers). The deep levels of the level path (thousands of small nodes, a few
// launches each) become this one launch.
constexpr int kMidStack = 16;  // pending nodes: the larger child waits, so at most log2(2048 / 65) + 1
constexpr int kMidWaves = kChunkBlock / kWave;

struct MidLds {
    uint4 stack[kMidStack];  // (begin, end, node id, prims | buffer << 31)
    uint32_t sbox[kMidStack][12];  // their centroid and full boxes (order-mapped)
    uint32_t cnt[kNumBuckets], pr[kNumBuckets], box[kNumBuckets][6];
    uint8_t bkt[kMidNode];  // bucket of each position of the node
};
// Resident grid: each wave takes nodes wave-stride from the list, whose length
// it reads on the device.
// Order-mapped min (k % 6 < 3) / max of a record's centroid and box into v.
__device__ __forceinline__ void add_rec_bounds(uint32_t v[12], const Rec &r) {
    const Box b = rec_box(r);
    const f3 c = box_centroid(b);
    v[0] = min(v[0], f2ord(c.x)); v[1] = min(v[1], f2ord(c.y)); v[2] = min(v[2], f2ord(c.z));
    v[3] = max(v[3], f2ord(c.x)); v[4] = max(v[4], f2ord(c.y)); v[5] = max(v[5], f2ord(c.z));
    v[6] = min(v[6], f2ord(b.lo.x)); v[7] = min(v[7], f2ord(b.lo.y)); v[8] = min(v[8], f2ord(b.lo.z));
    v[9] = max(v[9], f2ord(b.hi.x)); v[10] = max(v[10], f2ord(b.hi.y)); v[11] = max(v[11], f2ord(b.hi.z));
}

__global__ __launch_bounds__(kChunkBlock) void mid_top_kernel(BuildArgs A) {
    __shared__ MidLds s_lds[kMidWaves];
    MidLds &L = s_lds[threadIdx.x / kWave];
    const uint32_t lane = __lane_id();
    const uint32_t nitems = min(A.counters[kCntMidTop], A.cap_mid_top);
    for (uint32_t wid = blockIdx.x * kMidWaves + threadIdx.x / kWave; wid < nitems; wid += gridDim.x * kMidWaves) {
        {  // the list's node and its bounds; its descendants get theirs from the partition
            const WorkItem it = A.mid_top[wid];
            const Rec *R = pick(A.trec, it.buf);
            uint32_t v[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? 0xFFFFFFFFu : 0u;
            for (uint32_t p = it.begin + lane; p < it.end; p += kWave) add_rec_bounds(v, R[p]);
#pragma unroll
            for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? wave_min_u32_all(v[k]) : wave_max_u32_all(v[k]);
            if (lane == 0) {
                L.stack[0] = make_uint4(it.begin, it.end, it.node, it.prims | (it.buf << 31));
#pragma unroll
                for (int k = 0; k < 12; ++k) L.sbox[0][k] = v[k];
            }
        }
        int sp = 1;  // uniform
        wave_lds_sync();
        while (sp > 0) {
            --sp;
            const uint4 nd = L.stack[sp];
            const uint32_t begin = nd.x, end = nd.y, node = nd.z, prims = nd.w & 0x7FFFFFFFu, buf = nd.w >> 31;
            const uint32_t m = end - begin;
            const Rec *R = pick(A.trec, buf);
            Rec *W = pick(A.trec, 1 - buf);
            const Box cb = acc_box(L.sbox[sp]), fb = acc_box(L.sbox[sp] + 6);
            const int axis = box_max_dim(cb);
            if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) {  // the node stays unbuilt
                if (lane == 0) atomicOr(A.err, uint32_t(kErrDegenerateCentroids));
                continue;
            }
            const float base = sel3(cb.lo, axis);
            const float span = sel3(cb.hi, axis) - sel3(cb.lo, axis);
            if (lane < kNumBuckets) {
                L.cnt[lane] = 0u;
                L.pr[lane] = 0u;
#pragma unroll
                for (int j = 0; j < 6; ++j) L.box[lane][j] = j < 3 ? 0xFFFFFFFFu : 0u;
            }
            wave_lds_sync();
            // buckets, one row of 64 positions at a time (the next row's records
            // loaded ahead), merged into the node's statistics by LDS atomics (across
            // the wave reductions of same-bucket lanes cost more than the atomics
            // at these node sizes)
            Rec rn{};
            if (lane < m) rn = R[begin + lane];
            for (uint32_t p0 = 0; p0 < m; p0 += kWave) {
                const uint32_t p = p0 + lane;
                const Rec r = rn;
                if (p + kWave < m) rn = R[begin + p + kWave];
                if (p < m) {
                    const uint32_t k = bucket_of(sel3(box_centroid(rec_box(r)), axis), base, span);
                    L.bkt[p] = uint8_t(k);
                    atomicAdd(&L.cnt[k], 1u);
                    atomicAdd(&L.pr[k], r.b);
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        atomicMin(&L.box[k][d], f2ord(r.lo[d]));
                        atomicMax(&L.box[k][3 + d], f2ord(r.hi[d]));
                    }
                }
            }
            wave_lds_sync();
            // the price of the 23 splits (lane k = bucket k, and bucket 23 - k reversed)
            const bool in = lane < kNumBuckets;
            const uint32_t fl = in ? lane : 0u, rl = in ? kNumBuckets - 1 - lane : 0u;
            const uint32_t cnt = in ? L.pr[fl] : 0u, tcnt = in ? L.cnt[fl] : 0u, rcnt = in ? L.pr[rl] : 0u;
            uint32_t box[6], rbox[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {  // an untouched bucket holds the identities
                box[k] = in && L.cnt[fl] > 0 ? L.box[fl][k] : (k < 3 ? 0xFFFFFFFFu : 0u);
                rbox[k] = in && L.cnt[rl] > 0 ? L.box[rl][k] : (k < 3 ? 0xFFFFFFFFu : 0u);
            }
            const SplitPrice price = price_splits(cnt, tcnt, box, rcnt, rbox, box_area(fb));
            const uint32_t split = price.split, lp = price.left_prims, rp = price.right_prims;
            const uint32_t nl = price.left_treelets;  // left treelets
            // both end buckets hold a treelet, so neither side is empty (as top_decide_kernel)
            if (lp == 0u || rp == 0u || nl == 0u || nl == m) {
                if (lane == 0) atomicOr(A.err, uint32_t(kErrSplit));
                continue;
            }
            // stable partition into the other buffer, a row at a time (records loaded
            // ahead), with the two children's bounds
            uint32_t cv[2][12];
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int k = 0; k < 12; ++k) cv[c][k] = ((k % 6) < 3) ? 0xFFFFFFFFu : 0u;
            uint32_t lefts = 0;  // left positions in the rows before
            if (lane < m) rn = R[begin + lane];
            for (uint32_t p0 = 0; p0 < m; p0 += kWave) {
                const uint32_t p = p0 + lane;
                const Rec rec = rn;
                if (p + kWave < m) rn = R[begin + p + kWave];
                const bool left = p < m && L.bkt[p] <= split;
                const uint64_t bm = __ballot(left);
                if (p < m) {
                    const uint32_t lb = lefts + __popcll(bm & lanemask_lt());
                    if (left) W[begin + lb] = rec;
                    else W[begin + nl + (p - lb)] = rec;
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        if ((c == 0) == left) add_rec_bounds(cv[c], rec);
                }
                lefts += __popcll(bm);
            }
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int k = 0; k < 12; ++k)
                    cv[c][k] = ((k % 6) < 3) ? wave_min_u32_all(cv[c][k]) : wave_max_u32_all(cv[c][k]);
            // the children: lane 0 reads a single treelet back from W (written by this wave)
            __threadfence_block();
            if (lane == 0) {
                write_interior(A.nodes, node, fb, lp, uint32_t(axis));
                const uint32_t nb = 1 - buf;
                const WorkItem cl{begin, begin + nl, node + 1, lp, nb, {0, 0, 0}};
                const WorkItem cr{begin + nl, end, node + 2 * lp, prims - lp, nb, {0, 0, 0}};
                const bool right_big = m - nl > nl;
                int top = sp;
#pragma unroll
                for (int j = 0; j < 2; ++j) {  // the larger child waits below the smaller one
                    const bool take_right = (j == 0) == right_big;
                    const WorkItem cw = take_right ? cr : cl;
                    const uint32_t cs = cw.end - cw.begin;
                    if (cs == 1) {
                        emit_bottom_root(A, A.trec[nb][cw.begin], cw.node);
                    } else if (cs <= kSmallNode) {
                        push_item(A.small_top, A.counters + kCntSmallTop, A.cap_small_top, cw, A.err);
                    } else if (top < kMidStack) {
                        L.stack[top] = make_uint4(cw.begin, cw.end, cw.node, cw.prims | (nb << 31));
#pragma unroll
                        for (int k = 0; k < 12; ++k) L.sbox[top][k] = take_right ? cv[1][k] : cv[0][k];
                        ++top;
                    } else {
                        atomicOr(A.err, uint32_t(kErrCapacity));
                    }
                }
                sp = top;
            }
            sp = __shfl(sp, 0, kWave);
            wave_lds_sync();
        }
    }
}

// --------------------------------------------- wave-per-node subtree finisher

constexpr int kFinBlock = 256;
constexpr int kFinWaves = kFinBlock / kWave;

// Top-level nodes of at most this many treelets are split by one lane each
// (after the wave has split the larger ones cooperatively).
constexpr uint32_t kLaneNode = 4;

// Per-wave LDS working set of the finisher.
struct WaveLds {
    Rec rec[kSmallNode];
    Rec tmp[kSmallNode];
    uint4 stack[kSmallNode];
    uint32_t bkt[kSmallNode];
    float cost[kNumBuckets];
    uint2 leaf[kSmallNode];  // top level: (node id, wave-local index) of single-treelet leaves
    uint32_t acc[kNumBuckets][7];    // top level: bucket prim count and order-mapped box
    uint32_t leaf_node[kSmallNode];  // top level: node id of the single-treelet leaf at each position
    uint4 small[kSmallNode];         // top level: (lo, hi, node id) of nodes left to the lane phase
};

// Builds the whole subtree of one node with m0 <= 64 elements inside one wave.
// Elements sit one per lane in LDS; every split is a wave reduction plus a
// ballot-ranked stable partition. Top level (SAH over treelets): single-treelet
// leaves are recorded in L.leaf and their count returned. Bottom level
// (midpoint over primitives): leaves are written, with their primitive ids.
template <bool TOP>
__device__ uint32_t wave_build(const BuildArgs &A, const Rec *src, uint32_t begin, uint32_t m0, uint32_t node0,
                               uint32_t prims0, WaveLds &L) {
    const uint32_t lane = __lane_id();
    if (lane < m0) L.rec[lane] = src[begin + lane];
    if (lane == 0) L.stack[0] = make_uint4(0, m0, node0, prims0);
    int sp = 1;
    uint32_t nleaf = 0;
    wave_lds_sync();
    while (sp > 0) {
        --sp;
        const uint4 nd = L.stack[sp];
        const uint32_t lo = nd.x, hi = nd.y, id = nd.z, P = nd.w, m = hi - lo;
        const bool active = lane >= lo && lane < hi;
        Rec r;
        if (active) r = L.rec[lane];
        const Box b = active ? rec_box(r) : box_empty();
        const f3 c = box_centroid(b);
        uint32_t v[12];
        v[0] = active ? f2ord(c.x) : 0xFFFFFFFFu; v[1] = active ? f2ord(c.y) : 0xFFFFFFFFu;
        v[2] = active ? f2ord(c.z) : 0xFFFFFFFFu;
        v[3] = active ? f2ord(c.x) : 0u; v[4] = active ? f2ord(c.y) : 0u; v[5] = active ? f2ord(c.z) : 0u;
        v[6] = active ? f2ord(b.lo.x) : 0xFFFFFFFFu; v[7] = active ? f2ord(b.lo.y) : 0xFFFFFFFFu;
        v[8] = active ? f2ord(b.lo.z) : 0xFFFFFFFFu;
        v[9] = active ? f2ord(b.hi.x) : 0u; v[10] = active ? f2ord(b.hi.y) : 0u; v[11] = active ? f2ord(b.hi.z) : 0u;
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? wave_min_u32_all(v[k]) : wave_max_u32_all(v[k]);
        const Box cb = acc_box(v);
        const Box fb = acc_box(v + 6);
        const int axis = box_max_dim(cb);
        bool left;
        uint32_t lp = 0;
        if (TOP) {
            if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) {
                if (lane == 0) atomicOr(A.err, uint32_t(kErrDegenerateCentroids));
                continue;
            }
            const float base = sel3(cb.lo, axis);
            const float span = sel3(cb.hi, axis) - sel3(cb.lo, axis);
            const uint32_t bk = bucket_of(sel3(c, axis), base, span);
            if (active) L.bkt[lane] = bk;
            wave_lds_sync();
            const float total_sa = box_area(fb);
            // lane k prices the split after bucket k. From 16 elements on, lanes k and
            // k + 23 each take half of the elements and the halves are merged (exact
            // min / max and integer sums: the same boxes and counts)
            constexpr uint32_t kSplits = kNumBuckets - 1;
            const bool halves = m >= 16;
            const uint32_t kk = lane < kSplits ? lane : lane - kSplits;
            const bool pricing = lane < (halves ? 2 * kSplits : kSplits);
            const uint32_t mid = halves ? (lo + hi) / 2 : hi;
            const uint32_t j0 = lane < kSplits ? lo : mid, j1 = lane < kSplits ? mid : hi;
            Box bl = box_empty(), br = box_empty();
            uint32_t cl = 0, cr = 0;
            if (pricing)
                for (uint32_t j = j0; j < j1; ++j) {
                    const Rec q = L.rec[j];
                    if (L.bkt[j] <= kk) { bl = box_union(bl, rec_box(q)); cl += q.b; }
                    else { br = box_union(br, rec_box(q)); cr += q.b; }
                }
            if (halves) {  // all lanes: lane k takes lane k + 23's half
                const int src = int(lane + kSplits) & (kWave - 1);
                const Box ol{{__shfl(bl.lo.x, src, kWave), __shfl(bl.lo.y, src, kWave), __shfl(bl.lo.z, src, kWave)},
                             {__shfl(bl.hi.x, src, kWave), __shfl(bl.hi.y, src, kWave), __shfl(bl.hi.z, src, kWave)}};
                const Box orr{{__shfl(br.lo.x, src, kWave), __shfl(br.lo.y, src, kWave), __shfl(br.lo.z, src, kWave)},
                              {__shfl(br.hi.x, src, kWave), __shfl(br.hi.y, src, kWave), __shfl(br.hi.z, src, kWave)}};
                const uint32_t ocl = __shfl(cl, src, kWave), ocr = __shfl(cr, src, kWave);
                bl = box_union(bl, ol);
                br = box_union(br, orr);
                cl += ocl;
                cr += ocr;
            }
            if (lane < kSplits)
                L.cost[lane] = sah_cost((float(cl) * box_area(bl) + float(cr) * box_area(br)) / total_sa);
            wave_lds_sync();
            uint32_t split = 0;
            float best = L.cost[0];
            for (int k = 1; k < int(kNumBuckets) - 1; ++k) {
                const float ck = L.cost[k];
                if (ck < best) { best = ck; split = k; }
            }
            left = active && bk <= split;
            const uint64_t lm = __ballot(left);
            if (lm == 0 || __popcll(lm) == m) {  // midpoint fallback (hlbvh.cu:758-777)
                const float mid = sel3(box_centroid(cb), axis);
                left = active && sel3(c, axis) < mid;
            }
            lp = wave_sum_all(left ? r.b : 0u);
        } else {
            const float mid = sel3(box_centroid(cb), axis);
            left = active && sel3(c, axis) < mid;
        }
        const uint64_t lmask = __ballot(left);
        const uint64_t rmask = __ballot(active && !left);
        const uint32_t Lc = __popcll(lmask);
        if (Lc == 0 || Lc == m) {
            if (TOP) {
                if (lane == 0) atomicOr(A.err, uint32_t(kErrSplit));
                continue;
            }
            if (lane == 0) write_leaf(A.nodes, id, fb, begin + lo, m);  // all centroids on one side
            if (active) A.prim_order[begin + lane] = r.a;
            continue;
        }
        if (!TOP) lp = Lc;
        if (active) {
            const uint32_t dst = left ? lo + __popcll(lmask & lanemask_lt())
                                      : lo + Lc + __popcll(rmask & lanemask_lt());
            L.tmp[dst] = r;
        }
        wave_lds_sync();
        if (active) L.rec[lane] = L.tmp[lane];
        if (lane == 0) write_interior(A.nodes, id, fb, lp, axis);
        const uint4 ch[2] = {make_uint4(lo, lo + Lc, id + 1, lp), make_uint4(lo + Lc, hi, id + 2 * lp, P - lp)};
        wave_lds_sync();
        for (int k = 1; k >= 0; --k) {
            const uint4 cn = ch[k];
            if (cn.y - cn.x == 1) {
                if (TOP) {
                    if (lane == 0) L.leaf[nleaf] = make_uint2(cn.z, cn.x);
                    ++nleaf;
                } else if (lane == 0) {
                    const Rec q = L.rec[cn.x];
                    write_leaf(A.nodes, cn.z, rec_box(q), begin + cn.x, 1);
                    A.prim_order[begin + cn.x] = q.a;
                }
            } else {
                if (lane == 0) L.stack[sp] = cn;
                ++sp;
            }
        }
        wave_lds_sync();
    }
    return nleaf;
}

// Treelets of 2..16 primitives are built by lane_bottom_kernel, one lane per
// treelet, with the treelet's boxes held in registers: one batch of loads, then
// no memory reads. Treelets of up to kLaneSmall primitives and those of up to
// kLaneBuild go to two lists (the 8-slot variant needs half the registers).
constexpr uint32_t kLaneSmall = 8;
constexpr uint32_t kLaneBuild = 16;

__device__ __forceinline__ Rec load_rec(const Rec *p) {
    const float4 a = reinterpret_cast<const float4 *>(p)[0];
    const float4 b = reinterpret_cast<const float4 *>(p)[1];
    Rec r;
    r.lo[0] = a.x; r.lo[1] = a.y; r.lo[2] = a.z; r.a = __builtin_bit_cast(uint32_t, a.w);
    r.hi[0] = b.x; r.hi[1] = b.y; r.hi[2] = b.z; r.b = __builtin_bit_cast(uint32_t, b.w);
    return r;
}

__device__ __forceinline__ uint32_t nib(uint64_t w, uint32_t p) { return uint32_t(w >> (4 * p)) & 15u; }

// Box and primitive id of slot s, selected with compile-time indices only (the
// slot arrays stay in registers).
template <int MAXC>
__device__ __forceinline__ Box slot_box(const float (&lo)[MAXC][3], const float (&hi)[MAXC][3],
                                        const uint32_t (&pid)[MAXC], uint32_t s, uint32_t *id) {
    Box b{{lo[0][0], lo[0][1], lo[0][2]}, {hi[0][0], hi[0][1], hi[0][2]}};
    uint32_t q = pid[0];
#pragma unroll
    for (int k = 1; k < MAXC; ++k) {
        const bool m = s == uint32_t(k);
        b.lo.x = m ? lo[k][0] : b.lo.x; b.lo.y = m ? lo[k][1] : b.lo.y; b.lo.z = m ? lo[k][2] : b.lo.z;
        b.hi.x = m ? hi[k][0] : b.hi.x; b.hi.y = m ? hi[k][1] : b.hi.y; b.hi.z = m ? hi[k][2] : b.hi.z;
        q = m ? pid[k] : q;
    }
    *id = q;
    return b;
}

constexpr int kLaneBlock = 256;

// Midpoint subtree of one treelet (hlbvh.cu:828-857): at each node the centroid
// box's widest axis and its middle, a two-pointer partition on centroid < middle,
// and a leaf of all the node's primitives when one side is empty. Slot k holds the
// treelet's k-th primitive as loaded; `at` maps each position of the treelet's
// range to its slot (4 bits a position) and carries the partitions, so positions,
// leaf ranges and the primitive order are those of partitioning the records in
// place. Pending nodes (two primitives or more; a single one is written as its
// leaf at once) sit in four 64-bit registers as 16-bit entries (lo, hi, node
// offset); a treelet of c primitives keeps fewer than c of them.
template <int MAXC>
__global__ __launch_bounds__(kLaneBlock) __attribute__((amdgpu_waves_per_eu(MAXC <= 8 ? 3 : 2))) void lane_bottom_kernel(
    BuildArgs A, const WorkItem *items, const uint32_t *list_count, uint32_t cap) {
    const uint32_t nitems = min(*list_count, cap);  // the grid covers the list's capacity
    if (blockIdx.x * kLaneBlock >= nitems) return;  // the whole workgroup
    // The workgroup's treelets ordered by size (a counting sort in LDS), so that
    // each wave's lanes take similar numbers of steps.
    __shared__ uint32_t s_bin[MAXC + 2];
    __shared__ uint4 s_item[kLaneBlock];
    const uint32_t t = blockIdx.x * kLaneBlock + threadIdx.x;
    uint4 mine = make_uint4(0u, 0u, 0u, 0u);
    uint32_t key = MAXC + 1;  // past the end: no treelet
    if (t < nitems) {
        const WorkItem it = items[t];
        mine = make_uint4(it.begin, it.end - it.begin, it.node, 0u);
        key = min(mine.y, uint32_t(MAXC + 1));
    }
    if (threadIdx.x < MAXC + 2) s_bin[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t rank_in_bin = atomicAdd(&s_bin[key], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int k = 0; k < MAXC + 2; ++k) {
            const uint32_t c = s_bin[k];
            s_bin[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
    s_item[s_bin[key] + rank_in_bin] = mine;
    __syncthreads();
    const uint4 it = s_item[threadIdx.x];
    const uint32_t first = it.x, count = it.y, node0 = it.z;
    if (count == 0) return;  // past the list's end (sorted last)
    if (count < 2 || count > uint32_t(MAXC)) {  // the list holds 2..MAXC only
        atomicOr(A.err, uint32_t(kErrCapacity));
        return;
    }
    float lo[MAXC][3], hi[MAXC][3];
    uint32_t pid[MAXC];
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
        Rec r{};
        if (uint32_t(k) < count) r = load_rec(A.prec[0] + first + k);
        lo[k][0] = r.lo[0]; lo[k][1] = r.lo[1]; lo[k][2] = r.lo[2];
        hi[k][0] = r.hi[0]; hi[k][1] = r.hi[1]; hi[k][2] = r.hi[2];
        pid[k] = r.a;
    }
    uint64_t at = 0xFEDCBA9876543210ull;
    uint64_t st0 = uint64_t(0u | (count << 5)), st1 = 0, st2 = 0, st3 = 0;  // the root: [0, count), offset 0
    int sp = 1;
    while (sp > 0) {
        const uint32_t e = uint32_t(st0 & 0xFFFFu);
        st0 = (st0 >> 16) | (st1 << 48);
        st1 = (st1 >> 16) | (st2 << 48);
        st2 = (st2 >> 16) | (st3 << 48);
        st3 >>= 16;
        --sp;
        const uint32_t l = e & 31u, h = (e >> 5) & 31u, off = e >> 10;
        const uint32_t id = node0 + off;  // a node of >= 2 primitives
        uint32_t set = 0;
        for (uint32_t p = l; p < h; ++p) set |= 1u << nib(at, p);
        Box cb = box_empty(), fb = box_empty();
#pragma unroll
        for (int k = 0; k < MAXC; ++k) {
            if (!((set >> k) & 1u)) continue;
            const Box b{{lo[k][0], lo[k][1], lo[k][2]}, {hi[k][0], hi[k][1], hi[k][2]}};
            cb = box_add(cb, box_centroid(b));
            fb = box_union(fb, b);
        }
        const int axis = box_max_dim(cb);
        const float mid = sel3(box_centroid(cb), axis);
        uint32_t left = 0;
#pragma unroll
        for (int k = 0; k < MAXC; ++k) {
            const float c = sel3(box_centroid(Box{{lo[k][0], lo[k][1], lo[k][2]}, {hi[k][0], hi[k][1], hi[k][2]}}), axis);
            left |= (((set >> k) & 1u) && c < mid) ? (1u << k) : 0u;
        }
        // two-pointer partition of the positions [l, h)
        uint32_t i = l, j = h;
        while (i < j) {
            if ((left >> nib(at, i)) & 1u) {
                ++i;
            } else {
                --j;
                const uint64_t si = nib(at, i), sj = nib(at, j);
                at &= ~((15ull << (4 * i)) | (15ull << (4 * j)));
                at |= (sj << (4 * i)) | (si << (4 * j));
            }
        }
        const uint32_t L = i - l;
        if (L == 0 || L == h - l) {  // all centroids on one side: a multi-primitive leaf
            write_leaf(A.nodes, id, fb, first + l, h - l);
            for (uint32_t p = l; p < h; ++p) {
                uint32_t q;
                slot_box<MAXC>(lo, hi, pid, nib(at, p), &q);
                A.prim_order[first + p] = q;
            }
            continue;
        }
        write_interior(A.nodes, id, fb, L, axis);
        // a child of one primitive is its leaf now; the others are pushed, the
        // right one first (the left is taken next)
        const uint32_t cl[2] = {l, l + L}, ch[2] = {l + L, h}, co[2] = {off + 1, off + 2 * L};
#pragma unroll
        for (int c = 1; c >= 0; --c) {
            if (ch[c] - cl[c] == 1) {
                uint32_t q;
                const Box b = slot_box<MAXC>(lo, hi, pid, nib(at, cl[c]), &q);
                write_leaf(A.nodes, node0 + co[c], b, first + cl[c], 1);
                A.prim_order[first + cl[c]] = q;
            } else {
                st3 = (st3 << 16) | (st2 >> 48);
                st2 = (st2 << 16) | (st1 >> 48);
                st1 = (st1 << 16) | (st0 >> 48);
                st0 = (st0 << 16) | uint64_t(cl[c] | (ch[c] << 5) | (co[c] << 10));
                ++sp;
            }
        }
    }
}

// SAH subtree of one top-level node of 2..kLaneNode treelets, built by one lane
// over the wave's LDS records [lo, hi) (which no other lane touches). The split
// positions between two occupied buckets all give the same partition and the
// same cost, so the candidates are the occupied buckets but the last, in
// increasing order; the first strict minimum among them is the reference's
// first strict minimum over all 23 (hlbvh.cu:711-741). Each candidate's sides
// are unions over the node's elements (exact min / max, integer counts), the
// same boxes and counts as the reference's bucket unions. The partition is in
// place and unstable: the order of treelets inside a node changes no later
// decision. Pending nodes (at most three: the smaller child is taken next) sit
// in a register stack.
__device__ __forceinline__ void lane_sah_small(const BuildArgs &A, WaveLds &L, uint32_t lo, uint32_t hi, uint32_t id) {
    uint32_t st_r[4] = {0, 0, 0, 0}, st_id[4] = {0, 0, 0, 0};
    int sp = 0;
    for (;;) {
        const uint32_t m = hi - lo;
        if (m == 1) {
            L.leaf_node[lo] = id;
            if (sp == 0) return;
            --sp;
            lo = st_r[0] & 0xFFu;
            hi = st_r[0] >> 8;
            id = st_id[0];
#pragma unroll
            for (int k = 0; k < 3; ++k) { st_r[k] = st_r[k + 1]; st_id[k] = st_id[k + 1]; }
            continue;
        }
        uint32_t v[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? 0xFFFFFFFFu : 0u;
        for (uint32_t j = lo; j < hi; ++j) {
            const Box b = rec_box(L.rec[j]);
            const f3 c = box_centroid(b);
            v[0] = min(v[0], f2ord(c.x)); v[1] = min(v[1], f2ord(c.y)); v[2] = min(v[2], f2ord(c.z));
            v[3] = max(v[3], f2ord(c.x)); v[4] = max(v[4], f2ord(c.y)); v[5] = max(v[5], f2ord(c.z));
            v[6] = min(v[6], f2ord(b.lo.x)); v[7] = min(v[7], f2ord(b.lo.y)); v[8] = min(v[8], f2ord(b.lo.z));
            v[9] = max(v[9], f2ord(b.hi.x)); v[10] = max(v[10], f2ord(b.hi.y)); v[11] = max(v[11], f2ord(b.hi.z));
        }
        const Box cb = acc_box(v);
        const Box fb = acc_box(v + 6);
        const int axis = box_max_dim(cb);
        if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) {
            atomicOr(A.err, uint32_t(kErrDegenerateCentroids));
            return;
        }
        const float base = sel3(cb.lo, axis);
        const float span = sel3(cb.hi, axis) - sel3(cb.lo, axis);
        uint64_t bpack = 0;  // 5 bits per element
        uint32_t occ = 0;
        for (uint32_t j = lo; j < hi; ++j) {
            const uint32_t bk = bucket_of(sel3(box_centroid(rec_box(L.rec[j])), axis), base, span);
            bpack |= uint64_t(bk) << (5 * (j - lo));
            occ |= 1u << bk;
        }
        const float total_sa = box_area(fb);
        const uint32_t last = 31u - __builtin_clz(occ);
        uint32_t cand = occ & ~(1u << last);
        uint32_t split = 0, lp = 0;
        float best = 0.f;
        bool first = true;
        while (cand) {
            const uint32_t sb = __builtin_ctz(cand);
            cand &= cand - 1;
            Box bl = box_empty(), br = box_empty();
            uint32_t cl = 0, cr = 0;
            for (uint32_t j = lo; j < hi; ++j) {
                const Rec q = L.rec[j];
                if (uint32_t((bpack >> (5 * (j - lo))) & 31u) <= sb) { bl = box_union(bl, rec_box(q)); cl += q.b; }
                else { br = box_union(br, rec_box(q)); cr += q.b; }
            }
            const float cost = sah_cost((float(cl) * box_area(bl) + float(cr) * box_area(br)) / total_sa);
            if (first || cost < best) {
                best = cost;
                split = sb;
                lp = cl;
                first = false;
            }
        }
        // two-pointer partition on bucket <= split
        uint32_t i = lo, j = hi;
        while (i < j) {
            if (uint32_t((bpack >> (5 * (i - lo))) & 31u) <= split) {
                ++i;
            } else {
                --j;
                const Rec ri = L.rec[i];
                L.rec[i] = L.rec[j];
                L.rec[j] = ri;
                const uint64_t mi = 31ull << (5 * (i - lo)), mj = 31ull << (5 * (j - lo));
                const uint64_t bi = (bpack & mi) >> (5 * (i - lo)), bj = (bpack & mj) >> (5 * (j - lo));
                bpack = (bpack & ~(mi | mj)) | (bj << (5 * (i - lo))) | (bi << (5 * (j - lo)));
            }
        }
        const uint32_t nl = i - lo;
        if (nl == 0 || nl == m) {  // unreachable: the first and last buckets are both occupied
            atomicOr(A.err, uint32_t(kErrSplit));
            return;
        }
        write_interior(A.nodes, id, fb, lp, uint32_t(axis));
        // children: left [lo, lo + nl) as id + 1, right [lo + nl, hi) as id + 2 lp; the
        // smaller one next, the larger one pending
        const uint32_t llo = lo, lhi = lo + nl, lid = id + 1;
        const uint32_t rlo = lo + nl, rhi = hi, rid = id + 2 * lp;
        const bool left_first = nl <= m - nl;
        const uint32_t push_r = left_first ? (rlo | (rhi << 8)) : (llo | (lhi << 8));
        const uint32_t push_id = left_first ? rid : lid;
#pragma unroll
        for (int k = 3; k > 0; --k) { st_r[k] = st_r[k - 1]; st_id[k] = st_id[k - 1]; }
        st_r[0] = push_r;
        st_id[0] = push_id;
        ++sp;
        if (sp > 4) {  // cannot happen for <= 8 elements taken smaller side first
            atomicOr(A.err, uint32_t(kErrSplit));
            return;
        }
        lo = left_first ? llo : rlo;
        hi = left_first ? lhi : rhi;
        id = left_first ? lid : rid;
    }
}

// SAH subtree of one top-level node of m0 <= 64 treelets inside one wave
// (hlbvh.cu:649-826). Nodes larger than kLaneNode are split by the whole wave:
// element per lane, bounds by wave reductions, bucket statistics by LDS atomics
// (exact min / max of order-mapped floats, integer counts), then, as at the level
// path, lane k holds bucket k and inclusive prefix and suffix unions give it the
// cost of the split after bucket k (hlbvh.cu:711-731); the lowest lane holding
// the least cost is the first strict minimum (hlbvh.cu:734-741). Smaller nodes
// are collected and split one per lane (lane_sah_small). Every position ends
// with the node id of its single-treelet leaf in L.leaf_node.
__device__ __forceinline__ void wave_build_top(const BuildArgs &A, const Rec *src, uint32_t begin, uint32_t m0, uint32_t node0,
                               WaveLds &L) {
    const uint32_t lane = __lane_id();
    if (lane < m0) L.rec[lane] = src[begin + lane];
    if (lane == 0) L.stack[0] = make_uint4(0, m0, node0, 0);
    int sp = 1;
    uint32_t nsmall = 0;
    if (m0 <= kLaneNode) {  // straight to the lane phase
        sp = 0;
        if (lane == 0) L.small[0] = make_uint4(0, m0, node0, 0);
        nsmall = 1;
    }
    wave_lds_sync();
    while (sp > 0) {
        --sp;
        const uint4 nd = L.stack[sp];
        const uint32_t lo = nd.x, hi = nd.y, id = nd.z, m = hi - lo;
        const bool active = lane >= lo && lane < hi;
        Rec r;
        if (active) r = L.rec[lane];
        const Box b = active ? rec_box(r) : box_empty();
        const f3 c = box_centroid(b);
        uint32_t v[12];
        v[0] = active ? f2ord(c.x) : 0xFFFFFFFFu; v[1] = active ? f2ord(c.y) : 0xFFFFFFFFu;
        v[2] = active ? f2ord(c.z) : 0xFFFFFFFFu;
        v[3] = active ? f2ord(c.x) : 0u; v[4] = active ? f2ord(c.y) : 0u; v[5] = active ? f2ord(c.z) : 0u;
        v[6] = active ? f2ord(b.lo.x) : 0xFFFFFFFFu; v[7] = active ? f2ord(b.lo.y) : 0xFFFFFFFFu;
        v[8] = active ? f2ord(b.lo.z) : 0xFFFFFFFFu;
        v[9] = active ? f2ord(b.hi.x) : 0u; v[10] = active ? f2ord(b.hi.y) : 0u; v[11] = active ? f2ord(b.hi.z) : 0u;
#pragma unroll
        for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? wave_min_u32_all(v[k]) : wave_max_u32_all(v[k]);
        const Box cb = acc_box(v);
        const Box fb = acc_box(v + 6);
        const int axis = box_max_dim(cb);
        if (sel3(cb.lo, axis) == sel3(cb.hi, axis)) {
            if (lane == 0) atomicOr(A.err, uint32_t(kErrDegenerateCentroids));
            continue;
        }
        const float base = sel3(cb.lo, axis);
        const float span = sel3(cb.hi, axis) - sel3(cb.lo, axis);
        const uint32_t bk = bucket_of(sel3(c, axis), base, span);
        if (lane < kNumBuckets) {
            L.acc[lane][0] = 0u;
#pragma unroll
            for (int k = 1; k < 7; ++k) L.acc[lane][k] = k < 4 ? 0xFFFFFFFFu : 0u;
        }
        wave_lds_sync();
        if (active) {
            atomicAdd(&L.acc[bk][0], r.b);
            atomicMin(&L.acc[bk][1], f2ord(b.lo.x)); atomicMin(&L.acc[bk][2], f2ord(b.lo.y));
            atomicMin(&L.acc[bk][3], f2ord(b.lo.z));
            atomicMax(&L.acc[bk][4], f2ord(b.hi.x)); atomicMax(&L.acc[bk][5], f2ord(b.hi.y));
            atomicMax(&L.acc[bk][6], f2ord(b.hi.z));
        }
        wave_lds_sync();
        // lane k: bucket k; an empty bucket is the empty box
        const bool in = lane < kNumBuckets;
        const uint32_t fl = in ? lane : 0u, rl = in ? kNumBuckets - 1 - lane : 0u;  // the bucket, reversed
        const uint32_t cnt = in ? L.acc[fl][0] : 0u, rcnt = in ? L.acc[rl][0] : 0u;
        uint32_t box[6], rbox[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {  // an untouched bucket holds the identities
            box[k] = cnt > 0 ? L.acc[fl][1 + k] : (k < 3 ? 0xFFFFFFFFu : 0u);
            rbox[k] = rcnt > 0 ? L.acc[rl][1 + k] : (k < 3 ? 0xFFFFFFFFu : 0u);
        }
        const SplitPrice price = price_splits(cnt, 0u, box, rcnt, rbox, box_area(fb));
        const uint32_t split = price.split;
        uint32_t lp = price.left_prims;
        bool left = active && bk <= split;
        const uint64_t lm = __ballot(left);
        if (lm == 0 || __popcll(lm) == m) {  // midpoint fallback (hlbvh.cu:758-777)
            const float mid = sel3(box_centroid(cb), axis);
            left = active && sel3(c, axis) < mid;
            lp = wave_sum_all(left ? r.b : 0u);
        }
        const uint64_t lmask = __ballot(left);
        const uint64_t rmask = __ballot(active && !left);
        const uint32_t Lc = __popcll(lmask);
        if (Lc == 0 || Lc == m) {
            if (lane == 0) atomicOr(A.err, uint32_t(kErrSplit));
            continue;
        }
        if (active) {
            const uint32_t dst = left ? lo + __popcll(lmask & lanemask_lt())
                                      : lo + Lc + __popcll(rmask & lanemask_lt());
            L.tmp[dst] = r;
        }
        wave_lds_sync();
        if (active) L.rec[lane] = L.tmp[lane];
        if (lane == 0) write_interior(A.nodes, id, fb, lp, axis);
        const uint4 ch[2] = {make_uint4(lo, lo + Lc, id + 1, 0), make_uint4(lo + Lc, hi, id + 2 * lp, 0)};
        for (int k = 1; k >= 0; --k) {
            const uint4 cn = ch[k];
            const uint32_t cs = cn.y - cn.x;
            if (cs == 1) {
                if (lane == 0) L.leaf_node[cn.x] = cn.z;
            } else if (cs <= kLaneNode) {
                if (lane == 0) L.small[nsmall] = cn;
                ++nsmall;
            } else {
                if (lane == 0) L.stack[sp] = cn;
                ++sp;
            }
        }
        wave_lds_sync();
    }
    for (uint32_t k = lane; k < nsmall; k += kWave) {
        const uint4 nd = L.small[k];
        lane_sah_small(A, L, nd.x, nd.y, nd.z);
    }
    wave_lds_sync();
}

// Top-level finisher: one wave per node of <= 64 treelets. After the SAH
// subtree, each single-treelet leaf becomes the root of a midpoint subtree over
// its primitives: a leaf here when it has one, else queued by size for
// lane_bottom_kernel (<= 16), finish_bottom_kernel (<= 64) or the level path.
__global__ __launch_bounds__(kFinBlock) void finish_top_kernel(BuildArgs A, const WorkItem *items, uint32_t nitems) {
    __shared__ WaveLds s_lds[kFinWaves];
    const int w = threadIdx.x / kWave;
    const uint32_t lane = __lane_id();
    const uint32_t wid = blockIdx.x * kFinWaves + w;
    if (wid >= nitems) return;
    WaveLds &L = s_lds[w];
    const WorkItem it = items[wid];
    const uint32_t nleaf = it.end - it.begin;
    wave_build_top(A, pick(A.trec, it.buf), it.begin, nleaf, it.node, L);
    // lane k keeps the leaf at position k: a single primitive becomes the leaf;
    // larger treelets go to the bottom-level lists by size
    uint32_t my_node = 0, my_first = 0, my_count = 0;
    if (lane < nleaf) {
        const Rec t = L.rec[lane];
        my_node = L.leaf_node[lane];
        my_first = t.a;
        my_count = t.b;
        if (my_count == 1) {
            const Rec p = A.prec[0][my_first];
            write_leaf(A.nodes, my_node, rec_box(p), my_first, 1);
            A.prim_order[my_first] = p.a;
        } else if (my_count > kSmallNode) {
            push_item(A.large_bot[0], A.counters + kCntLargeBot0, A.cap_large_bot,
                      WorkItem{my_first, my_first + my_count, my_node, my_count, 0, {0, 0, 0}}, A.err);
        }
    }
    const WorkItem tw{my_first, my_first + my_count, my_node, my_count, 0, {0, 0, 0}};
    const bool real = lane < nleaf && my_count >= 2;
    wave_push(real && my_count <= kLaneSmall, tw, A.lane_bot[0], A.counters + kCntLaneBot0, A.cap_lane_bot[0], A.err);
    wave_push(real && my_count > kLaneSmall && my_count <= kLaneBuild, tw, A.lane_bot[1], A.counters + kCntLaneBot1,
          A.cap_lane_bot[1], A.err);
    wave_push(real && my_count > kLaneBuild && my_count <= kSmallNode, tw, A.small_bot, A.counters + kCntSmallBot,
          A.cap_small_bot, A.err);
}

// Bottom-level finisher for nodes of <= 64 primitives left by the level path.
__global__ __launch_bounds__(kFinBlock) void finish_bottom_kernel(BuildArgs A, const WorkItem *items, uint32_t nitems) {
    __shared__ WaveLds s_lds[kFinWaves];
    const int w = threadIdx.x / kWave;
    const uint32_t wid = blockIdx.x * kFinWaves + w;
    if (wid >= nitems) return;
    const WorkItem it = items[wid];
    wave_build<false>(A, pick(A.prec, it.buf), it.begin, it.end - it.begin, it.node, it.prims, s_lds[w]);
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void zero_counters_kernel(uint32_t *c, int n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0u;
}

__global__ void init_cbounds_kernel(uint32_t *cb) {
    if (threadIdx.x < 6) cb[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
}

}  // namespace

// ------------------------------------------------------------------ host side

#define PBRTM_TRY(x)                         \
    do {                                     \
        hipError_t _e = (x);                 \
        if (_e != hipSuccess) return _e;     \
    } while (0)

static hipError_t read_counters(BuildScratch &S, hipStream_t stream) {
    PBRTM_TRY(hipMemcpyAsync(S.host_counters, S.counters.ptr, kNumCounters * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, stream));
    return hipStreamSynchronize(stream);
}

// Levels enqueued between two reads of the item count. Grids between reads use
// upper bounds: a level has at most twice the previous level's large items
// (each splits in two), at most elements / (kSmallNode + 1) of them (they are
// disjoint and larger than kSmallNode), and at most elements / kChunk + items
// chunk tasks.
constexpr int kLevelsPerSync = 4;

template <bool TOP>
static hipError_t run_levels(BuildArgs &A, BuildScratch &S, hipStream_t stream, RecPair recs, int cnt0,
                             int cnt1, WorkItem *list0, WorkItem *list1, uint32_t elements, uint32_t list_cap,
                             uint32_t *levels) {
    WorkItem *cur = list0, *nxt = list1;
    int ccur = cnt0, cnxt = cnt1;
    uint32_t *const counters = S.counters.as<uint32_t>();
    uint32_t *const ntasks = counters + kCntTasks;
    const uint32_t item_cap = std::min(list_cap, elements / (kSmallNode + 1) + 1);
    PBRTM_TRY(read_counters(S, stream));
    uint32_t known = S.host_counters[ccur];  // exact item count of the next level
    uint32_t enqueued = 0;
    while (known > 0) {
        // bounds for the next kLevelsPerSync levels, and scratch sized for them
        const uint32_t nb_max = uint32_t(std::min<uint64_t>(uint64_t(known) << (kLevelsPerSync - 1), item_cap));
        const uint32_t tb_max = elements / kChunk + nb_max + 1;
        PBRTM_TRY(S.acc.reserve(size_t(nb_max) * sizeof(Acc)));
        PBRTM_TRY(S.splits.reserve(size_t(nb_max) * sizeof(Split)));
        PBRTM_TRY(S.task_off.reserve(size_t(nb_max + 1) * sizeof(uint32_t)));
        PBRTM_TRY(S.tasks.reserve(size_t(tb_max) * sizeof(uint2)));
        PBRTM_TRY(S.task_left.reserve(size_t(tb_max + 1) * sizeof(uint32_t)));
        PBRTM_TRY(S.task_lprims.reserve(size_t(tb_max + 1) * sizeof(uint32_t)));
        PBRTM_TRY(S.scan_scratch.reserve(size_t(tb_max / 4096 + 2) * sizeof(uint32_t)));
        uint32_t *task_off = S.task_off.as<uint32_t>();
        uint2 *tasks = S.tasks.as<uint2>();
        uint32_t *tleft = S.task_left.as<uint32_t>();
        uint32_t *tlp = S.task_lprims.as<uint32_t>();
        Acc *acc = S.acc.as<Acc>();
        Split *splits = S.splits.as<Split>();
        uint32_t nb = known;
        for (int k = 0; k < kLevelsPerSync; ++k) {
            if (++enqueued > 100000) return hipErrorUnknown;  // every level strictly shrinks its nodes
            const uint32_t tb = elements / kChunk + nb + 1;
            const uint32_t *ncur = counters + ccur;
            level_setup_kernel<<<uint32_t((uint64_t(nb) * kAccWords + 255) / 256), 256, 0, stream>>>(
                cur, ncur, nb, task_off, acc, counters + cnxt);
            PBRTM_TRY(exclusive_scan_u32(task_off, task_off, nb, ntasks, S.scan_scratch, stream));
            task_fill_kernel<<<(tb + 255) / 256, 256, 0, stream>>>(ncur, nb, task_off, ntasks, tasks);
            chunk_reduce_kernel<<<tb, kChunkBlock, 0, stream>>>(tasks, ntasks, cur, recs, acc);
            if (TOP) chunk_buckets_kernel<<<tb, kChunkBlock, 0, stream>>>(tasks, ntasks, cur, recs, acc, A.err);
            decide_kernel<TOP><<<TOP ? nb : (nb + 63) / 64, 64, 0, stream>>>(cur, ncur, acc, splits);
            chunk_count_kernel<TOP><<<tb, kChunkBlock, 0, stream>>>(tasks, ntasks, cur, recs, splits, tleft, tlp);
            PBRTM_TRY(exclusive_scan_u32(tleft, tleft, tb, tleft + tb, S.scan_scratch, stream));
            if (TOP) PBRTM_TRY(exclusive_scan_u32(tlp, tlp, tb, tlp + tb, S.scan_scratch, stream));
            chunk_scatter_kernel<TOP><<<tb, kChunkBlock, 0, stream>>>(tasks, ntasks, cur, recs, splits, task_off,
                                                                      tleft);
            children_kernel<TOP><<<(nb + 63) / 64, 64, 0, stream>>>(A, cur, ncur, splits, acc, task_off, tleft, tlp,
                                                                    nxt, counters + cnxt,
                                                                    counters + (TOP ? kCntLevelsTop : kCntLevelsBot));
            PBRTM_TRY(hipGetLastError());
            std::swap(cur, nxt);
            std::swap(ccur, cnxt);
            nb = std::min(2 * nb, item_cap);
        }
        PBRTM_TRY(read_counters(S, stream));
        known = S.host_counters[ccur];
    }
    *levels = S.host_counters[TOP ? kCntLevelsTop : kCntLevelsBot];
    return hipSuccess;
}


// The fused top level: three launches per level, levels enqueued kLevelsPerSync
// at a time between reads of the node count (grids cover upper bounds; the
// kernels read the level's counts on the device).
static hipError_t run_top_levels(BuildArgs &A, BuildScratch &S, hipStream_t stream, RecPair recs,
                                 uint32_t num_treelets, uint32_t n, uint32_t *levels) {
    uint32_t *const counters = S.counters.as<uint32_t>();
    const uint32_t cap_items = num_treelets / (kSmallNode + 1) + 2;
    const uint32_t cap_tasks = num_treelets / kChunk + cap_items + 1;
    PBRTM_TRY(S.top_items.reserve(2 * size_t(cap_items) * sizeof(TopItem)));
    PBRTM_TRY(S.top_acc.reserve(2 * size_t(cap_items) * sizeof(TopAcc)));
    PBRTM_TRY(S.top_tasks.reserve(2 * size_t(cap_tasks) * sizeof(uint2)));
    PBRTM_TRY(S.top_hist.reserve(size_t(cap_tasks) * kNumBuckets * sizeof(uint32_t)));
    PBRTM_TRY(S.top_left.reserve(size_t(cap_tasks) * sizeof(uint32_t)));
    PBRTM_TRY(S.top_bkt.reserve(size_t(num_treelets)));
    TopLevel TL{};
    TL.items[0] = S.top_items.as<TopItem>();
    TL.items[1] = TL.items[0] + cap_items;
    TL.acc[0] = S.top_acc.as<TopAcc>();
    TL.acc[1] = TL.acc[0] + cap_items;
    TL.tasks[0] = S.top_tasks.as<uint2>();
    TL.tasks[1] = TL.tasks[0] + cap_tasks;
    TL.hist = S.top_hist.as<uint32_t>();
    TL.left = S.top_left.as<uint32_t>();
    TL.bkt = S.top_bkt.as<uint8_t>();
    TL.cap_items = cap_items;
    TL.cap_tasks = cap_tasks;
    TL.num = num_treelets;
    top_root_kernel<<<1, 256, 0, stream>>>(TL, num_treelets, n, counters);
    const uint32_t rgrid = std::min<uint32_t>((num_treelets + kChunkBlock - 1) / kChunkBlock, 1024u);
    top_root_box_kernel<<<rgrid, kChunkBlock, 0, stream>>>(A.trec[0], num_treelets, TL);
    PBRTM_TRY(hipGetLastError());
    uint32_t known = 1, p = 0, enqueued = 0;
    const uint32_t item_cap = cap_items;
    // the first batch covers the fewest levels that can bring every node down to
    // kMidNode treelets (halving at best), so that few count reads follow
    int batch = kLevelsPerSync;
    for (uint64_t c = kMidNode; c < num_treelets; c <<= 1) ++batch;
    batch = std::max(kLevelsPerSync, batch - 3);
    while (known > 0) {
        uint32_t nb = known;
        for (int k = 0; k < batch; ++k) {
            if (++enqueued > 100000) return hipErrorUnknown;  // every level strictly shrinks its nodes
            const uint32_t tb = num_treelets / kChunk + nb + 1;
            top_bucket_kernel<<<tb, kChunkBlock, 0, stream>>>(TL, p, recs, counters, A.err);
            top_decide_kernel<<<nb, 64, 0, stream>>>(TL, p, counters, A.err);
            top_scatter_kernel<<<tb, kChunkBlock, 0, stream>>>(TL, p, recs, counters);
            top_children_kernel<<<(nb + 63) / 64, 64, 0, stream>>>(A, TL, p, counters + kCntLevelsTop);
            top_acc_reset_kernel<<<1024, 256, 0, stream>>>(TL, 1 - p, counters);
            PBRTM_TRY(hipGetLastError());
            p = 1 - p;
            nb = std::min(2 * nb, item_cap);
        }
        batch = kLevelsPerSync;
        PBRTM_TRY(read_counters(S, stream));
        known = S.host_counters[kCntTopItems0 + p];
    }
    *levels = S.host_counters[kCntLevelsTop];
    return hipSuccess;
}

hipError_t morton_sort_device(const float *d_tri, uint32_t n, BuildScratch &S, uint32_t **keys_out,
                              uint32_t **vals_out, hipStream_t stream) {
    PBRTM_TRY(S.counters.reserve(kNumCounters * sizeof(uint32_t)));
    for (int k = 0; k < 2; ++k) {
        PBRTM_TRY(S.keys[k].reserve(size_t(n) * 4));
        PBRTM_TRY(S.vals[k].reserve(size_t(n) * 4));
    }
    PBRTM_TRY(S.hist.reserve((kSortPasses * kSortRadix + 64) * 4));
    uint32_t *counters = S.counters.as<uint32_t>();
    uint32_t *hist = S.hist.as<uint32_t>();
    uint32_t *cbounds = hist + kSortPasses * kSortRadix;
    const uint32_t grid_n = (n + 255) / 256;
    zero_counters_kernel<<<1, 64, 0, stream>>>(counters, kNumCounters);
    *keys_out = S.keys[0].as<uint32_t>();
    *vals_out = S.vals[0].as<uint32_t>();
    if (n == 0) return hipGetLastError();
    fill_u32_kernel<<<12, 256, 0, stream>>>(hist, kSortPasses * kSortRadix, 0u);
    init_cbounds_kernel<<<1, 64, 0, stream>>>(cbounds);
    centroid_bounds_kernel<<<min(grid_n, 1024u), 256, 0, stream>>>(d_tri, n, cbounds, counters + kCntErr);
    morton_kernel<<<min(grid_n, 512u), 256, 0, stream>>>(d_tri, n, cbounds, S.keys[0].as<uint32_t>(),
                                                          S.vals[0].as<uint32_t>(), hist);
    return onesweep_sort_30(S.keys[0].as<uint32_t>(), S.vals[0].as<uint32_t>(), S.keys[1].as<uint32_t>(),
                            S.vals[1].as<uint32_t>(), n, hist, counters + kCntErr, S.sort, stream, keys_out,
                            vals_out);
}

hipError_t build_hlbvh_device(const float *d_tri, uint32_t n, BuildScratch &S, BvhDevice &out,
                              uint32_t *device_error, hipStream_t stream, BuildPhaseTimes *times) {
    out.n = n;
    out.num_treelets = out.top_levels = out.bottom_levels = out.small_top = out.small_bottom = 0;
    *device_error = 0;
    if (n == 0) return hipSuccess;
    PBRTM_TRY(S.counters.reserve(kNumCounters * sizeof(uint32_t)));
    if (!S.host_counters) PBRTM_TRY(hipHostMalloc((void **)&S.host_counters, kNumCounters * sizeof(uint32_t)));
    for (int k = 0; k < 2; ++k) {
        PBRTM_TRY(S.keys[k].reserve(size_t(n) * 4));
        PBRTM_TRY(S.vals[k].reserve(size_t(n) * 4));
        PBRTM_TRY(S.prec[k].reserve(size_t(n) * sizeof(Rec)));
    }
    PBRTM_TRY(S.hist.reserve((kSortPasses * kSortRadix + 64) * 4));
    PBRTM_TRY(out.nodes.reserve(size_t(2 * uint64_t(n) - 1) * sizeof(BNode)));
    PBRTM_TRY(out.prim_order.reserve(size_t(n) * 4));
    PBRTM_TRY(S.tri_sorted.reserve(size_t(n) * 36));
    uint32_t *counters = S.counters.as<uint32_t>();
    uint32_t *hist = S.hist.as<uint32_t>();
    uint32_t *cbounds = hist + kSortPasses * kSortRadix;
    const uint32_t grid_n = (n + 255) / 256;

    if (times) hipEventRecord(times->ev[0], stream);
    uint32_t *K = nullptr, *V = nullptr;
    PBRTM_TRY(morton_sort_device(d_tri, n, S, &K, &V, stream));
    if (times) hipEventRecord(times->ev[1], stream);
    uint32_t *flags = (K == S.keys[0].as<uint32_t>()) ? S.keys[1].as<uint32_t>() : S.keys[0].as<uint32_t>();
    uint32_t *tidx = (V == S.vals[0].as<uint32_t>()) ? S.vals[1].as<uint32_t>() : S.vals[0].as<uint32_t>();
    Rec *prec0 = S.prec[0].as<Rec>();
    S.sorted_pid = V;
    gather_kernel<<<grid_n, 256, 0, stream>>>(d_tri, n, K, V, prec0, S.tri_sorted.as<float>(), flags);
    PBRTM_TRY(exclusive_scan_u32(flags, tidx, n, counters + kCntTreelets, S.scan_scratch, stream));
    PBRTM_TRY(read_counters(S, stream));
    const uint32_t D = S.host_counters[kCntTreelets];
    out.num_treelets = D;
    PBRTM_TRY(S.trec[0].reserve(size_t(D) * sizeof(Rec)));
    PBRTM_TRY(S.trec[1].reserve(size_t(D) * sizeof(Rec)));
    uint32_t *tstart = K;  // sorted keys are dead once flags exist
    treelet_start_kernel<<<grid_n, 256, 0, stream>>>(flags, tidx, n, tstart);
    treelet_rec_kernel<<<(D + 255) / 256, 256, 0, stream>>>(tstart, D, n, prec0, S.trec[0].as<Rec>());
    if (times) hipEventRecord(times->ev[2], stream);

    BuildArgs A{};
    A.nodes = out.nodes.as<BNode>();
    A.prim_order = out.prim_order.as<uint32_t>();
    A.trec[0] = S.trec[0].as<Rec>();
    A.trec[1] = S.trec[1].as<Rec>();
    A.prec[0] = prec0;
    A.prec[1] = S.prec[1].as<Rec>();
    A.cap_small_top = D / 2 + 64;
    A.cap_large_top = D / 64 + 64;
    A.cap_small_bot = n / 2 + 64;
    A.cap_large_bot = n / 64 + 64;
    // disjoint treelets of >= 2 and >= 9 primitives
    A.cap_lane_bot[0] = std::min(D, n / 2) + 64;
    A.cap_lane_bot[1] = std::min(D, n / (kLaneSmall + 1)) + 64;
    A.cap_mid_top = D / (kSmallNode + 1) + 64;  // disjoint nodes of > 64 treelets
    const size_t list_items = size_t(A.cap_small_top) + A.cap_small_bot + 2 * size_t(A.cap_large_top) +
                              2 * size_t(A.cap_large_bot) + A.cap_lane_bot[0] + A.cap_lane_bot[1] +
                              A.cap_mid_top;
    PBRTM_TRY(S.lists.reserve(list_items * sizeof(WorkItem)));
    WorkItem *lp = S.lists.as<WorkItem>();
    A.small_top = lp; lp += A.cap_small_top;
    A.small_bot = lp; lp += A.cap_small_bot;
    A.large_top[0] = lp; lp += A.cap_large_top;
    A.large_top[1] = lp; lp += A.cap_large_top;
    A.large_bot[0] = lp; lp += A.cap_large_bot;
    A.large_bot[1] = lp; lp += A.cap_large_bot;
    A.lane_bot[0] = lp; lp += A.cap_lane_bot[0];
    A.lane_bot[1] = lp; lp += A.cap_lane_bot[1];
    A.mid_top = lp;
    A.counters = counters;
    A.err = counters + kCntErr;
    const RecPair trecs{{A.trec[0], A.trec[1]}}, precs{{A.prec[0], A.prec[1]}};

    if (D > kMidNode) {  // the fused top level down to nodes of <= 2048 treelets
        PBRTM_TRY(run_top_levels(A, S, stream, trecs, D, n, &out.top_levels));
    } else {
        root_kernel<<<1, 64, 0, stream>>>(A, D, n);
    }
    // the nodes of 65..2048 treelets: a resident grid, the list's length read on
    // the device (bounded by what the list can hold)
    static int mid_grid = 0;
    if (!mid_grid) {
        int per_cu = 0, cus = 0, dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        mid_grid = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mid_top_kernel, kChunkBlock, 0) ==
                           hipSuccess && per_cu > 0 ? per_cu * cus : 4 * cus;
    }
    if (D > kSmallNode)
        mid_top_kernel<<<std::min<uint32_t>(mid_grid, (A.cap_mid_top + kMidWaves - 1) / kMidWaves), kChunkBlock, 0,
                         stream>>>(A);
    // then the nodes of <= 64, a wave each (an exact grid lets the hardware balance their uneven work)
    PBRTM_TRY(read_counters(S, stream));
    out.small_top = std::min(S.host_counters[kCntSmallTop], A.cap_small_top);
    if (out.small_top > 0)
        finish_top_kernel<<<(out.small_top + kFinWaves - 1) / kFinWaves, kFinBlock, 0, stream>>>(
            A, A.small_top, out.small_top);
    if (times) hipEventRecord(times->ev[3], stream);
    // one-lane treelet subtrees: grids over the lists' capacities, counts read on the device
    lane_bottom_kernel<kLaneSmall><<<(A.cap_lane_bot[0] + kLaneBlock - 1) / kLaneBlock, kLaneBlock, 0, stream>>>(
        A, A.lane_bot[0], counters + kCntLaneBot0, A.cap_lane_bot[0]);
    lane_bottom_kernel<kLaneBuild><<<(A.cap_lane_bot[1] + kLaneBlock - 1) / kLaneBlock, kLaneBlock, 0, stream>>>(
        A, A.lane_bot[1], counters + kCntLaneBot1, A.cap_lane_bot[1]);
    // run_levels ends on a read of the counters after its last level: the small
    // bottom list is complete in it
    PBRTM_TRY(run_levels<false>(A, S, stream, precs, kCntLargeBot0, kCntLargeBot1, A.large_bot[0],
                                A.large_bot[1], n, A.cap_large_bot, &out.bottom_levels));
    out.small_bottom = std::min(S.host_counters[kCntSmallBot], A.cap_small_bot);
    if (out.small_bottom > 0)
        finish_bottom_kernel<<<(out.small_bottom + kFinWaves - 1) / kFinWaves, kFinBlock, 0, stream>>>(
            A, A.small_bot, out.small_bottom);
    if (times) hipEventRecord(times->ev[4], stream);
    PBRTM_TRY(hipGetLastError());
    PBRTM_TRY(read_counters(S, stream));
    *device_error = S.host_counters[kCntErr];
    if (times) hipEventRecord(times->ev[5], stream);
    return hipSuccess;
}

hipError_t leaf_order(const BuildScratch &S, BvhDevice &bvh, float *tri_leaf, hipStream_t stream) {
    if (bvh.n == 0) return hipSuccess;
    leaf_order_kernel<<<(bvh.n + 255) / 256, 256, 0, stream>>>(S.sorted_pid, S.tri_sorted.as<float>(), bvh.n,
                                                               bvh.prim_order.as<uint32_t>(), tri_leaf);
    return hipGetLastError();
}

void BuildScratch::release() {
    DevBuf *all[] = {&counters, &hist, &keys[0], &keys[1], &vals[0], &vals[1], &prec[0], &prec[1],
                     &trec[0], &trec[1], &lists, &acc, &splits, &tasks, &task_off, &task_left,
                     &task_lprims, &scan_scratch, &sort.status, &sort.offsets, &sort.counters,
                     &top_items, &top_acc, &top_tasks, &top_hist, &top_left, &top_bkt, &tri_sorted};
    for (DevBuf *b : all) b->release();
    if (host_counters) hipHostFree(host_counters);
    host_counters = nullptr;
}

}  // namespace pbrtm
