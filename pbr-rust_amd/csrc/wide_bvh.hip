// 8-wide BVH: collapse of the binary build tree and the closest/any-hit
// traversal over it.
//
// Collapse: each wide node stands for one binary interior node; its frontier
// is grown by repeatedly opening the frontier interior node of largest surface
// area until 8 slots are filled (the greedy collapse of Wald et al. 2008 /
// Ylitie et al. 2017). Child boxes are the binary nodes' exact FP32 boxes,
// quantized outwards to 8 bits per plane (one 128-byte node).
//
// Traversal keeps the reference's semantics exactly (accelerator/hlbvh.cu:179-289):
// for every ray octant a wide node stores the order in which the binary
// depth-first, near-child-first traversal reaches its slots, and a slot is
// entered only if its box passes the slab test against the current best t at
// the moment the binary traversal would pop it. Box tests are monotone under
// box inclusion, so skipping the collapsed inner binary boxes changes no
// decision: the sequence of triangle tests, and therefore every hit (ties
// included), equals the binary traversal's. Quantized boxes only contain the
// exact ones, so they add box entries and triangle tests the binary traversal
// would not make. Those can only find a hit the reference misses when the hit
// lies on the exact leaf box's boundary (a tie with the current best t on a box
// face, or a rounding of the entry distance): the first hit in a leaf is
// therefore kept only if the leaf's exact box (the union of its primitives'
// boxes) passes the slab test with the t at which the leaf was entered, exactly
// as the binary traversal tests it. Each visit costs one 128-byte node instead
// of up to seven dependent 32-byte fetches.
#include "pbrtm_geometry.h"
#include "wave_ops.h"
#include "pbrtm_shapes.h"
#include "wide_bvh.h"

#include <algorithm>

namespace pbrtm {

namespace {

struct Job {
    uint32_t bnode, wnode, depth, prims;  // prims: primitives under the binary node
};

// 2^e for e in [-126, 127], exactly.
__device__ __forceinline__ float pow2f(int e) { return __uint_as_float(uint32_t(e + 127) << 23); }

// Encodes n (<= 8) child boxes into w's origin, exponents and quantized planes;
// unused slots get the inverted box (lo 255, hi 0). ulo / uhi: the union of the
// child boxes (the collapsed binary node's own box); box(s, lo, hi) fetches child
// s. Steps are powers of two, so the divisions below are exact products.
template <typename BoxOf>
__device__ __forceinline__ void encode_node(WNode *w, const float ulo[3], const float uhi[3], BoxOf box, int n) {
    float sc[3], inv[3];
    uint32_t ex = 0u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ext = uhi[a] - ulo[a];
        int e = -126;
        if (ext > 0.f) {  // smallest power-of-two step with ext / step <= 253
            int e2;
            frexpf(ext / 253.0f, &e2);
            e = e2;
            if (ldexpf(1.0f, e2 - 1) * 253.0f >= ext) e = e2 - 1;
        }
        e = max(-126, min(126, e));
        sc[a] = pow2f(e);
        inv[a] = pow2f(-e);
        ex |= (uint32_t(e) & 0xFFu) << (8 * a);
    }
    uint64_t plo[3] = {0, 0, 0}, phi[3] = {0, 0, 0};
#pragma unroll
    for (int s = 0; s < kWide; ++s) {
        float l[3] = {0.f, 0.f, 0.f}, h[3] = {0.f, 0.f, 0.f};
        if (s < n) box(s, l, h);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int ql = 255, qh = 0;
            if (s < n) {
                ql = int(fminf(255.f, fmaxf(0.f, floorf((l[a] - ulo[a]) * inv[a]))));
                while (ql > 0 && fmaf(float(ql), sc[a], ulo[a]) > l[a]) --ql;
                qh = int(fminf(255.f, fmaxf(0.f, ceilf((h[a] - ulo[a]) * inv[a]))));
                while (qh < 255 && fmaf(float(qh), sc[a], ulo[a]) < h[a]) ++qh;
            }
            plo[a] |= uint64_t(ql) << (8 * s);
            phi[a] |= uint64_t(qh) << (8 * s);
        }
    }
    uint4 *dst = reinterpret_cast<uint4 *>(w);
    dst[0] = make_uint4(__float_as_uint(ulo[0]), __float_as_uint(ulo[1]), __float_as_uint(ulo[2]), ex);
    dst[1] = make_uint4(uint32_t(plo[0]), uint32_t(plo[0] >> 32), uint32_t(plo[1]), uint32_t(plo[1] >> 32));
    dst[2] = make_uint4(uint32_t(plo[2]), uint32_t(plo[2] >> 32), uint32_t(phi[0]), uint32_t(phi[0] >> 32));
    dst[3] = make_uint4(uint32_t(phi[1]), uint32_t(phi[1] >> 32), uint32_t(phi[2]), uint32_t(phi[2] >> 32));
}

__device__ __forceinline__ Box bnode_box(const BNode &b) {
    return {{b.lo[0], b.lo[1], b.lo[2]}, {b.hi[0], b.hi[1], b.hi[2]}};
}
__device__ __forceinline__ bool bnode_leaf(const BNode &b) { return (b.a & kLeafFlag) != 0; }

// Collapse counters: [4] levels that had jobs, [kCollapseJobs + L % 3] the job
// count J(L) of level L and [kCollapseBase + L % 3] the wide index B(L) of its
// first job (three slots each, in rotation). Level L's jobs are the wide nodes
// B(L) .. B(L) + J(L) - 1, so the children it appends get B(L) + J(L) + their job
// index at level L + 1, and B(L) + J(L) of the first level without jobs is the
// number of wide nodes.
constexpr uint32_t kCollapseJobs = 8;
constexpr uint32_t kCollapseBase = 12;

__global__ void wide_root_kernel(const BNode *bn, uint32_t n, Job *jobs, uint32_t *counters, WNode *wn) {
    if (threadIdx.x != 0) return;
    const BNode r = bn[0];
    if (bnode_leaf(r)) {  // the whole tree is one leaf: a wide root with one leaf slot
        encode_node(wn, r.lo, r.hi, [&](int, float *l, float *h) {
            for (int a = 0; a < 3; ++a) l[a] = r.lo[a], h[a] = r.hi[a];
        }, 1);
        for (int s = 0; s < kWide; ++s) {
            const bool use = s == 0;
            wn[0].ref[s][0] = use ? r.a : kWideEmpty;
            wn[0].ref[s][1] = use ? r.b : 0u;
        }
        counters[kCollapseJobs] = 0;  // no jobs at level 0
        counters[kCollapseBase] = 1;  // one node, written here
        return;
    }
    jobs[0] = Job{0, 0, 1, n};
    counters[kCollapseJobs] = 1;
    counters[kCollapseBase] = 0;
}

constexpr int kCollapseBlock = 64;
constexpr int kCollapseWaves = kCollapseBlock / kWave;
// A binary subtree of P primitives holds its 2P - 1 nodes contiguously in
// preorder (left child = id + 1, right child = id + 2 * P_left). Up to
// kPrefetchPrims primitives (31 nodes, at most 9 lines of 128 bytes), a job
// loads every line of its subtree at once, so that the dependent node fetches
// of the collapse find them in cache instead of waiting on memory one by one.
constexpr uint32_t kPrefetchPrims = 16;
constexpr uint32_t kPrefetchLines = 9;

// One wide node per lane. Frontier state lives in registers: every array below
// is indexed only with compile-time indices (fully unrolled loops with selects).
// Every thread of the workgroup (one wave) takes part, live or not: the child
// reservation is one atomic per workgroup.
__device__ __forceinline__ void collapse_one(const BNode *__restrict__ bn, const Job *__restrict__ jobs, uint32_t i,
                                             bool live, Job *__restrict__ next, uint32_t *__restrict__ next_count,
                                             uint32_t node_base, uint32_t *s_wave, uint32_t &s_base,
                                             WNode *__restrict__ wn, uint32_t *__restrict__ order, uint32_t cap,
                                             uint32_t num_bnodes, uint32_t *__restrict__ err) {
    const Job job = live ? jobs[i] : Job{0, 0, 0, 0};
    uint32_t touch = 0u;
    if (live && job.prims >= 2u && job.prims <= kPrefetchPrims) {
        const uint32_t *words = reinterpret_cast<const uint32_t *>(bn);
        const uint32_t last_node = min(job.bnode + 2u * job.prims - 2u, num_bnodes - 1u);
        const uint32_t first_line = job.bnode >> 2, last_line = last_node >> 2;
#pragma unroll
        for (uint32_t k = 1; k < kPrefetchLines; ++k)  // line 0 holds the root, loaded next
            if (first_line + k <= last_line) touch ^= words[size_t(first_line + k) * 32u];
    }
    uint32_t fid[kWide];      // frontier binary ids, left to right
    uint32_t fp[kWide];       // their primitive counts
    float farea[kWide];       // surface area, or -1 for a leaf / unused slot
    uint32_t fwa[kWide], fwb[kWide];  // the frontier nodes' a and b words (children, or leaf range)
    uint32_t oid[kWide - 1], ol[kWide - 1], orr[kWide - 1], oax[kWide - 1];  // opened nodes
#pragma unroll
    for (int k = 0; k < kWide; ++k) { fid[k] = 0u; fp[k] = 0u; farea[k] = -1.0f; fwa[k] = fwb[k] = 0u; }
#pragma unroll
    for (int k = 0; k < kWide - 1; ++k) { oid[k] = 0xFFFFFFFFu; ol[k] = orr[k] = oax[k] = 0u; }
    int nF = 0, nE = 0;
    if (live) {
        const BNode r = bn[job.bnode];
        oid[0] = job.bnode; ol[0] = r.a; orr[0] = r.b & 0x3FFFFFFFu; oax[0] = r.b >> 30;
        nE = 1;
        const BNode c0 = bn[ol[0]], c1 = bn[orr[0]];
        fid[0] = ol[0]; farea[0] = bnode_leaf(c0) ? -1.0f : box_area(bnode_box(c0));
        fid[1] = orr[0]; farea[1] = bnode_leaf(c1) ? -1.0f : box_area(bnode_box(c1));
        fp[0] = (orr[0] - job.bnode) >> 1;
        fp[1] = job.prims - fp[0];
        fwa[0] = c0.a; fwb[0] = c0.b;
        fwa[1] = c1.a; fwb[1] = c1.b;
        nF = 2;
        while (nF < kWide) {  // open the frontier interior node of largest surface area
            int best = -1;
            float best_area = -1.0f;
#pragma unroll
            for (int k = 0; k < kWide; ++k)
                if (farea[k] > best_area) { best_area = farea[k]; best = k; }
            if (best < 0) break;
            // the opened node's child ids are already in registers: one dependent
            // fetch (its two children, independent of each other) per step
            uint32_t bid = 0, ba = 0, bb = 0, bp = 0;
#pragma unroll
            for (int k = 0; k < kWide; ++k) {
                bid = (k == best) ? fid[k] : bid;
                ba = (k == best) ? fwa[k] : ba;
                bb = (k == best) ? fwb[k] : bb;
                bp = (k == best) ? fp[k] : bp;
            }
            const uint32_t cl = ba, cr = bb & 0x3FFFFFFFu;
            const uint32_t pl = (cr - bid) >> 1, pr = bp - pl;
            const BNode nl = bn[cl], nr = bn[cr];
#pragma unroll
            for (int k = 0; k < kWide - 1; ++k) {
                const bool here = k == nE;
                oid[k] = here ? bid : oid[k];
                ol[k] = here ? cl : ol[k];
                orr[k] = here ? cr : orr[k];
                oax[k] = here ? (bb >> 30) : oax[k];
            }
            ++nE;
            const float al = bnode_leaf(nl) ? -1.0f : box_area(bnode_box(nl));
            const float ar = bnode_leaf(nr) ? -1.0f : box_area(bnode_box(nr));
#pragma unroll
            for (int k = kWide - 1; k >= 0; --k) {  // shift right of best by one, then place the children
                const uint32_t prev_id = k > 0 ? fid[k - 1] : 0u, prev_p = k > 0 ? fp[k - 1] : 0u;
                const float prev_a = k > 0 ? farea[k - 1] : -1.0f;
                const uint32_t prev_wa = k > 0 ? fwa[k - 1] : 0u, prev_wb = k > 0 ? fwb[k - 1] : 0u;
                if (k > best + 1) { fid[k] = prev_id; fp[k] = prev_p; farea[k] = prev_a; fwa[k] = prev_wa; fwb[k] = prev_wb; }
                else if (k == best + 1) { fid[k] = cr; fp[k] = pr; farea[k] = ar; fwa[k] = nr.a; fwb[k] = nr.b; }
                else if (k == best) { fid[k] = cl; fp[k] = pl; farea[k] = al; fwa[k] = nl.a; fwb[k] = nl.b; }
            }
            ++nF;
        }
    }
    // interior children need job slots (and so wide node indices): one atomic per
    // workgroup. s_wave is read only between the two barriers and written only
    // after the second, so consecutive calls may share it.
    uint32_t n_int = 0;
#pragma unroll
    for (int s = 0; s < kWide; ++s) n_int += (s < nF && farea[s] >= 0.0f) ? 1u : 0u;
    const uint32_t incl = wave_inclusive_add(n_int);
    const uint32_t wv = threadIdx.x / kWave;
    if (__lane_id() == kWave - 1) s_wave[wv] = incl;
    __syncthreads();
    uint32_t wave_off = 0, total = 0;
#pragma unroll
    for (uint32_t k = 0; k < uint32_t(kCollapseWaves); ++k) {
        const uint32_t t = s_wave[k];
        wave_off += k < wv ? t : 0u;
        total += t;
    }
    if (threadIdx.x == 0) s_base = total > 0 ? atomicAdd(next_count, total) : 0u;
    __syncthreads();
    const uint32_t jbase = s_base + wave_off + incl - n_int;
    const uint32_t base = node_base + jbase;
    if (!live) return;
    if (base + n_int > cap) {
        atomicOr(err, uint32_t(kErrCapacity));
        return;
    }
    uint32_t refw[kWide][2];
    uint32_t k_int = 0;
#pragma unroll
    for (int s = 0; s < kWide; ++s) {
        if (s < nF) {
            if (farea[s] >= 0.0f) {
                const uint32_t idx = base + k_int;
                next[jbase + k_int] = Job{fid[s], idx, job.depth + 1, fp[s]};
                ++k_int;
                refw[s][0] = idx;
                refw[s][1] = 0u;
            } else {
                refw[s][0] = fwa[s];  // kLeafFlag | first (same bit as kWideLeaf)
                refw[s][1] = fwb[s];
            }
        } else {
            refw[s][0] = kWideEmpty;
            refw[s][1] = 0u;
        }
    }
    WNode *w = wn + job.wnode;
    uint4 *wr = reinterpret_cast<uint4 *>(w->ref);
#pragma unroll
    for (int k = 0; k < 4; ++k) wr[k] = make_uint4(refw[2 * k][0], refw[2 * k][1], refw[2 * k + 1][0], refw[2 * k + 1][1]);
    // depth-first, near-first slot order per octant (the binary traversal's order).
    // Codes: 0..7 = frontier slot, 8..14 = opened node; the DFS stack is eight
    // 4-bit codes packed in one register.
    // per opened node k: its left and right codes in byte k of lr (left in the low
    // nibble) and its split axis in bits 2k..2k+1 of axes
    uint64_t lr = 0;
    uint32_t axes = 0;
#pragma unroll
    for (int k = 0; k < kWide - 1; ++k) {
        uint32_t lc = 0xFu, rc = 0xFu;
#pragma unroll
        for (int j = 0; j < kWide - 1; ++j) {
            if (j < nE && oid[j] == ol[k]) lc = 8u + j;
            if (j < nE && oid[j] == orr[k]) rc = 8u + j;
        }
#pragma unroll
        for (int s = 0; s < kWide; ++s) {
            if (s < nF && fid[s] == ol[k]) lc = s;
            if (s < nF && fid[s] == orr[k]) rc = s;
        }
        lr |= uint64_t(lc | (rc << 4)) << (8 * k);
        axes |= (oax[k] & 3u) << (2 * k);
    }
    uint32_t *ow = order + size_t(job.wnode) * 8;
#pragma unroll 1
    for (uint32_t oct = 0; oct < 8; ++oct) {
        uint32_t st = 8u;  // root = opened node 0
        int sp = 1;
        uint32_t pm = 0xCCCCCCCCu;  // unused positions: 12 (a byte-permute selector that yields 0)
        int p = 0;
        while (sp > 0) {
            const uint32_t c = st & 0xFu;
            st >>= 4;
            --sp;
            if (c >= 8u) {
                const uint32_t e = uint32_t(lr >> (8 * (c - 8u)));
                const uint32_t l = e & 0xFu, r = (e >> 4) & 0xFu, ax = (axes >> (2 * (c - 8u))) & 3u;
                const bool neg = (oct >> ax) & 1u;
                st = (st << 8) | (neg ? r : l) | ((neg ? l : r) << 4);  // near on top
                sp += 2;
            } else {
                pm = (pm & ~(0xFu << (4 * p))) | (c << (4 * p));
                ++p;
            }
        }
        ow[oct] = pm;
    }
    // child boxes re-read from the binary nodes (cached), all eight fetches issued
    // together (after the order words, whose registers are then free); their union
    // is the collapsed node's own box
    float cbl[kWide][3], cbh[kWide][3];
    const BNode root = bn[job.bnode];
#pragma unroll
    for (int s = 0; s < kWide; ++s) {
        const BNode c = bn[s < nF ? fid[s] : fid[0]];
#pragma unroll
        for (int a = 0; a < 3; ++a) { cbl[s][a] = c.lo[a]; cbh[s][a] = c.hi[a]; }
    }
    encode_node(w, root.lo, root.hi, [&](int s, float *l, float *h) {
        for (int a = 0; a < 3; ++a) l[a] = cbl[s][a], h[a] = cbh[s][a];
    }, nF);
    // the prefetched words are consumed here (a depth of all ones never occurs)
    if (touch == 0x9E3779B9u && job.depth == 0xFFFFFFFFu) atomicOr(err, 0u);
}

// One level of the collapse. Job counts rotate over three counter slots: level L
// reads its count from slot L % 3, appends its children's jobs to slot (L + 1) % 3
// and clears slot (L + 2) % 3 (level L - 1's, read by no one any more; level L + 1
// appends to it), so no kernel runs between levels; B(L + 1) = B(L) + J(L) goes
// to the base slot (L + 1) % 3 the same way. A grid of resident workgroups walks
// the jobs kCollapseBlock at a time; a level that had jobs is counted (counters[4]).
__global__ __launch_bounds__(kCollapseBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void wide_collapse_kernel(
    const BNode *__restrict__ bn, const Job *__restrict__ jobs, Job *__restrict__ next,
    uint32_t *__restrict__ counters, WNode *__restrict__ wn, uint32_t *__restrict__ order, uint32_t cap,
    uint32_t num_bnodes, uint32_t *__restrict__ err, uint32_t level) {
    __shared__ uint32_t s_wave[kCollapseWaves];
    __shared__ uint32_t s_base;
    const uint32_t njobs = counters[kCollapseJobs + level % 3];
    const uint32_t node_base = counters[kCollapseBase + level % 3] + njobs;  // B(L + 1)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[kCollapseJobs + (level + 2) % 3] = 0u;
        counters[kCollapseBase + (level + 1) % 3] = node_base;
        if (njobs > 0) counters[4] += 1u;
    }
    uint32_t *next_count = counters + kCollapseJobs + (level + 1) % 3;
    const uint32_t stride = gridDim.x * kCollapseBlock;
    for (uint32_t base = blockIdx.x * kCollapseBlock; base < njobs; base += stride) {
        const uint32_t i = base + threadIdx.x;
        collapse_one(bn, jobs, i, i < njobs, next, next_count, node_base, s_wave, s_base, wn, order, cap,
                     num_bnodes, err);
    }
}

// Entry distance of one child box, or +inf when the slab test fails for a
// reason that does not depend on the ray's current t_max (bounds3.h:146-195).
// box_hit(..., t) == (entry < t). n / f: the near and far plane of each axis,
// already chosen by the ray's direction signs.
__device__ __forceinline__ float box_entry_nf(float nx, float ny, float nz, float fx, float fy, float fz, f3 o,
                                              f3 inv) {
    float t_min = (nx - o.x) * inv.x;
    float t_max = widen_far((fx - o.x) * inv.x);
    const float ty_min = (ny - o.y) * inv.y;
    const float ty_max = widen_far((fy - o.y) * inv.y);
    if (t_min > ty_max || ty_min > t_max) return __builtin_inff();
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    const float tz_min = (nz - o.z) * inv.z;
    const float tz_max = widen_far((fz - o.z) * inv.z);
    if (t_min > tz_max || tz_min > t_max) return __builtin_inff();
    if (tz_min > t_min) t_min = tz_min;
    if (tz_max < t_max) t_max = tz_max;
    return t_max > 0.0f ? t_min : __builtin_inff();
}

// The same for a box given by its lower and upper planes.
__device__ __forceinline__ float box_entry(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv,
                                           const int neg[3]) {
    return box_entry_nf(neg[0] ? hx : lx, neg[1] ? hy : ly, neg[2] ? hz : lz, neg[0] ? lx : hx, neg[1] ? ly : hy,
                        neg[2] ? lz : hz, o, inv);
}

constexpr int kWideBlock = 256;

// Ray queue compaction: active rays (t_max > 0) are appended to `queue` with
// one atomic per workgroup (wave ballots + a workgroup prefix), inactive ones
// get their miss written here. The traversal then runs on full waves.
// Active rays (t_max > 0) into a dense queue, in ray order within each tile of
// kCompactBlock * kCompactItems rays; inactive rays report a miss. One queue
// reservation (atomic) per tile.
constexpr uint32_t kCompactBlock = 1024, kCompactItems = 4;
__global__ __launch_bounds__(kCompactBlock) void compact_rays_kernel(const float4 *__restrict__ rays, uint32_t n,
                                                                     uint32_t *__restrict__ queue,
                                                                     uint32_t *__restrict__ count,
                                                                     uint32_t *__restrict__ prim_out,
                                                                     float *__restrict__ t_out,
                                                                     float *__restrict__ bary_out) {
    __shared__ uint32_t s_waves[kCompactBlock / kWave];
    __shared__ uint32_t s_base;
    const uint32_t tile = blockIdx.x * kCompactBlock * kCompactItems;
    bool act[kCompactItems];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCompactItems; ++k) {
        const uint32_t i = tile + k * kCompactBlock + threadIdx.x;
        const bool in = i < n;
        act[k] = in && rays[2 * size_t(i)].w > 0.f;
        if (in && !act[k]) {
            prim_out[i] = kNoHit;
            if (t_out) t_out[i] = 0.f;
            if (bary_out) {
                bary_out[3 * size_t(i)] = 0.f;
                bary_out[3 * size_t(i) + 1] = 0.f;
                bary_out[3 * size_t(i) + 2] = 0.f;
            }
        }
        mine += act[k] ? 1u : 0u;
    }
    uint32_t tot;
    block_exclusive_add<kCompactBlock>(mine, s_waves, &tot);
    if (threadIdx.x == 0) s_base = tot ? atomicAdd(count, tot) : 0u;
    __syncthreads();
    uint32_t run = s_base;
#pragma unroll
    for (uint32_t k = 0; k < kCompactItems; ++k) {
        uint32_t part;
        const uint32_t pre = block_exclusive_add<kCompactBlock>(act[k] ? 1u : 0u, s_waves, &part);
        if (act[k]) queue[run + pre] = tile + k * kCompactBlock + threadIdx.x;
        run += part;
    }
}

// Traversal state of one ray.
struct WideRay {
    f3 o, d, inv;
    float best_t;
    uint32_t best, oct, node, pmask, sp;
    TriHit hit;
};

// The tree one ray is traced against.
struct WideTree {
    const WNode *nodes;
    const uint32_t *order;
    const float *tri_leaf;
    const uint32_t *prim_order;
    const pt::DShape *shapes;
    uint32_t num_nodes, num_prims;
};

__device__ __forceinline__ WideTree tree_of(const WideTrace &B) {
    return WideTree{B.nodes, B.order, B.tri_leaf, B.prim_order, B.shapes, B.num_nodes, B.num_prims};
}

// The exact box of the binary leaf at positions [first, first + cnt) (the union of
// its primitives' boxes: the leaf-ordered vertices, or an analytic primitive's
// box stored as the corners lo, hi, lo) against the ray with t_max (bounds3.h:146-195).
__device__ __forceinline__ bool leaf_box_hit(const WideTree &T, uint32_t first, uint32_t cnt,
                                                        const WideRay &R, float t_max) {
    Box b = box_empty();
    for (uint32_t k = first; k < first + cnt; ++k) {
        const float4 *t4 = reinterpret_cast<const float4 *>(T.tri_leaf) + 3ull * k;
        const float4 q0 = t4[0], q1 = t4[1], q2 = t4[2];
        b = box_union(b, tri_box(f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, f3{q1.z, q1.w, q2.x}));
    }
    const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
    const int neg[3] = {int(R.oct & 1u), int((R.oct >> 1) & 1u), int((R.oct >> 2) & 1u)};
    return box_hit(lo, hi, R.o, R.inv, neg, t_max);
}

// Loads ray i; true when it has an interval to trace (t_max > 0).
__device__ __forceinline__ bool wide_ray_start(WideRay &R, const float4 *__restrict__ rays, uint32_t i) {
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    R.o = {r0.x, r0.y, r0.z};
    R.d = {r1.x, r1.y, r1.z};
    R.best_t = r0.w;
    R.best = kNoHit;
    R.hit = TriHit{0.f, 0.f, 0.f, 0.f};
    R.inv = {1.0f / R.d.x, 1.0f / R.d.y, 1.0f / R.d.z};
    R.oct = uint32_t(R.inv.x < 0.f) | (uint32_t(R.inv.y < 0.f) << 1) | (uint32_t(R.inv.z < 0.f) << 2);
    R.node = 0;
    R.pmask = 0xFFu;
    R.sp = 0;
    return R.best_t > 0.f;
}

__device__ __forceinline__ void wide_ray_finish(const WideRay &R, uint32_t i, uint32_t *__restrict__ prim_out,
                                                float *__restrict__ t_out, float *__restrict__ bary_out) {
    prim_out[i] = R.best;
    if (t_out) t_out[i] = R.best == kNoHit ? 0.f : R.hit.t;
    if (bary_out) {
        bary_out[3 * i] = R.hit.b0;
        bary_out[3 * i + 1] = R.hit.b1;
        bary_out[3 * i + 2] = R.hit.b2;
    }
}

// One node visit of one ray: tests the node's eight child boxes, then takes
// the node's pending children in depth-first position order (leaves are
// intersected on the spot, the first interior child is descended into, the
// rest of the node is pushed), or pops the stack. Returns false once the ray
// is finished. qi: the ray's queue slot, which owns the ray's spill area.
template <bool COUNT, bool SHAPES>
__device__ __forceinline__ bool wide_step(WideRay &R, const WideTree &T, const WideTrace &B, int any_hit,
                                          uint32_t qi, uint32_t *__restrict__ err,
                                          uint2 (*s_stack)[kWideBlock], uint32_t &visits, uint32_t &boxes,
                                          uint32_t &tests) {
    if (R.node >= T.num_nodes) {  // only a corrupt layout gets here
        atomicOr(err, uint32_t(kErrStack));
        return false;
    }
    const int neg[3] = {int(R.oct & 1u), int((R.oct >> 1) & 1u), int((R.oct >> 2) & 1u)};
    const WNode *W = T.nodes + R.node;
    const uint32_t perm = T.order[R.node * 8 + R.oct];
    if (COUNT) { ++visits; boxes += kWide; }
    // all eight child boxes (4 x 16-byte loads: origin, exponents, 48 plane
    // bytes), decoded to FP32 planes, hit mask in slot order
    const uint4 *q = reinterpret_cast<const uint4 *>(W);
    const uint4 hd = q[0];
    const float org[3] = {__uint_as_float(hd.x), __uint_as_float(hd.y), __uint_as_float(hd.z)};
    const float scl[3] = {pow2f(int(int8_t(hd.w & 0xFFu))), pow2f(int(int8_t((hd.w >> 8) & 0xFFu))),
                          pow2f(int(int8_t((hd.w >> 16) & 0xFFu)))};
    uint32_t qw[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint4 v = q[1 + k];
        qw[4 * k] = v.x;
        qw[4 * k + 1] = v.y;
        qw[4 * k + 2] = v.z;
        qw[4 * k + 3] = v.w;
    }
    // near / far plane bytes of each axis (plane p: bytes 8p..8p+7, words 2p and
    // 2p + 1), chosen once per node by the ray's direction signs
    // (bitwise blends: a select between array elements would become an indexed read)
    uint32_t nw[3][2], fw[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t m = 0u - uint32_t(neg[a]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t lo = qw[2 * a + h], hi = qw[2 * (a + 3) + h];
            nw[a][h] = (hi & m) | (lo & ~m);
            fw[a][h] = (lo & m) | (hi & ~m);
        }
    }
    uint32_t hs = 0;
#pragma unroll
    for (int s = 0; s < kWide; ++s) {
        float np[3], fp[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            np[a] = fmaf(float((nw[a][s / 4] >> (8 * (s % 4))) & 0xFFu), scl[a], org[a]);
            fp[a] = fmaf(float((fw[a][s / 4] >> (8 * (s % 4))) & 0xFFu), scl[a], org[a]);
        }
        const float e = box_entry_nf(np[0], np[1], np[2], fp[0], fp[1], fp[2], R.o, R.inv);
        hs |= (e < R.best_t) ? (1u << s) : 0u;
    }
    // ... then in depth-first position order: hm bit p = hs bit (slot at position p).
    // The hit bits become bytes (slot s -> byte s of {d1, d0}), a byte permute
    // gathers them by the position's slot code (an unused position's code 12
    // yields 0), and a multiply packs each 4 bytes back into 4 bits.
    const uint32_t d0 = ((hs & 0xFu) * 0x00204081u) & 0x01010101u;
    const uint32_t d1 = ((hs >> 4) * 0x00204081u) & 0x01010101u;
    const uint32_t plo = perm & 0xFFFFu, phi = perm >> 16;
    const uint32_t sel0 = (plo & 0xFu) | ((plo & 0xF0u) << 4) | ((plo & 0xF00u) << 8) | ((plo & 0xF000u) << 12);
    const uint32_t sel1 = (phi & 0xFu) | ((phi & 0xF0u) << 4) | ((phi & 0xF00u) << 8) | ((phi & 0xF000u) << 12);
    const uint32_t g0 = __builtin_amdgcn_perm(d1, d0, sel0), g1 = __builtin_amdgcn_perm(d1, d0, sel1);
    const uint32_t hm = (((g0 * 0x01020408u) >> 24) & 0xFu) | ((((g1 * 0x01020408u) >> 24) & 0xFu) << 4);
    uint32_t pmask = R.pmask & hm;
    const float t_at_visit = R.best_t;
    while (pmask) {
        const uint32_t p = __builtin_ctz(pmask);
        pmask &= pmask - 1;
        const uint32_t s = (perm >> (4 * p)) & 0xFu;
        if (R.best_t != t_at_visit) {  // a leaf of this node shortened the ray: re-test at pop time
            if (COUNT) ++boxes;
            float bx[6];
#pragma unroll
            for (int pp = 0; pp < 6; ++pp) bx[pp] = fmaf(float(W->q[pp][s]), scl[pp % 3], org[pp % 3]);
            const float e = box_entry(bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], R.o, R.inv, neg);
            if (!(e < R.best_t)) continue;
        }
        const uint32_t ref0 = W->ref[s][0];
        if (ref0 & kWideLeaf) {
            const uint32_t first = ref0 & ~kWideLeaf, cnt = W->ref[s][1];
            if (uint64_t(first) + cnt > T.num_prims) {
                atomicOr(err, uint32_t(kErrStack));
                return false;
            }
            const float t_leaf = R.best_t;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const float4 *t4 = reinterpret_cast<const float4 *>(T.tri_leaf) + 3ull * k;
                const float4 q0 = t4[0], q1 = t4[1], q2 = t4[2];
                if (COUNT) ++tests;
                TriHit h;
                bool hit, robust = false;
                if (SHAPES && __float_as_uint(q2.y) != 0u) {
                    pt::QHit qh;
                    hit = pt::shape_hit(T.shapes, __float_as_uint(q2.y) - 1u, R.o, R.d, R.best_t, &qh);
                    h = TriHit{qh.p.x, qh.p.y, qh.p.z, qh.t};
                } else {
                    hit = tri_hit(R.o, R.d, R.best_t, f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y},
                                  f3{q1.z, q1.w, q2.x}, &h, &robust);
                }
                if (hit) {
                    // The first hit in this leaf: would the binary traversal have entered
                    // the leaf? Its exact box (the union of its primitives' boxes) against
                    // the t at entry; the quantized slot box only contains it. A robust
                    // triangle hit (tri_hit: clear of the edges and of t_max by its error
                    // bounds, where t_max is the t at entry) passes every box holding the
                    // triangle; only the others (ties on box faces, edge and grazing hits,
                    // analytic shapes) assemble the leaf's box from its primitives.
                    if (R.best_t == t_leaf && !robust && !leaf_box_hit(T, first, cnt, R, t_leaf)) break;
                    R.best = T.prim_order[k];
                    R.hit = h;
                    R.best_t = h.t;
                    if (any_hit) return false;
                }
            }
            continue;
        }
        if (pmask) {  // the rest of this node waits
            const uint32_t sp = R.sp;
            if (sp < kLdsStack) {
                s_stack[sp][threadIdx.x] = make_uint2(R.node, pmask);
            } else if (sp - kLdsStack < B.spill_per_ray) {
                B.spill[uint64_t(qi) * B.spill_per_ray + (sp - kLdsStack)] = make_uint2(R.node, pmask);
            } else {
                atomicOr(err, uint32_t(kErrStack));
                return false;
            }
            R.sp = sp + 1;
        }
        R.node = ref0;
        R.pmask = 0xFFu;
        return true;
    }
    if (R.sp == 0) return false;
    const uint32_t sp = --R.sp;
    const uint2 e = sp < kLdsStack ? s_stack[sp][threadIdx.x]
                                   : B.spill[uint64_t(qi) * B.spill_per_ray + (sp - kLdsStack)];
    R.node = e.x;
    R.pmask = e.y;
    return true;
}

template <bool COUNT>
__device__ __forceinline__ void add_counts(unsigned long long *counts, uint32_t visits, uint32_t boxes,
                                           uint32_t tests) {
    if (!COUNT) return;
    const unsigned long long v = wave_sum(visits), b = wave_sum(boxes), tt = wave_sum(tests);
    if (__lane_id() == 0) {
        atomicAdd(&counts[0], v);
        atomicAdd(&counts[1], b);
        atomicAdd(&counts[2], tt);
    }
}

// Persistent traversal: a grid that fills the GPU once; every wave walks its own
// sequence of queue slots (row r of the queue, wave w: slots r * G * 64 + w * 64 +
// lane, G = waves in the grid) and a lane whose ray has finished takes the wave's
// next slot as soon as kRefill lanes are idle, so that the waves keep most lanes
// busy instead of waiting for their slowest ray. Each ray is traced on its own
// (wide_ray_start, wide_step until it reports the ray finished, wide_ray_finish),
// so which lane or wave takes it changes nothing. MULTI: every ray selects its own
// tree, jw[ray_job[i >> job_shift]] (jobs of a batched render); the spill area
// stays the launch's.
constexpr uint32_t kRefill = 8;

template <bool COUNT, bool MULTI, bool SHAPES>
__device__ __forceinline__ void trace_wide_persistent_body(const WideTrace &B, const float4 *__restrict__ rays,
                                                           const uint32_t *__restrict__ queue,
                                                           const uint32_t *__restrict__ queue_len, int any_hit,
                                                           uint32_t *__restrict__ prim_out,
                                                           float *__restrict__ t_out, float *__restrict__ bary_out,
                                                           uint32_t *__restrict__ err,
                                                           unsigned long long *__restrict__ counts,
                                                           const WideTrace *__restrict__ jw,
                                                           const uint32_t *__restrict__ ray_job, uint32_t job_shift,
                                                           uint2 (*s_stack)[kWideBlock]) {
    const uint32_t n = *queue_len;
    const uint32_t lane = __lane_id();
    const uint64_t stride = uint64_t(gridDim.x) * kWideBlock;  // queue slots per row
    const uint64_t base = uint64_t(blockIdx.x) * kWideBlock + (threadIdx.x & ~uint32_t(kWave - 1));
    WideTree T = tree_of(B);  // MULTI: the tree of the lane's current ray
    uint32_t visits = 0, boxes = 0, tests = 0;
    WideRay R;
    bool live = false;
    uint32_t qi = 0, i = 0;
    uint32_t next = 0;  // wave-uniform: positions of the wave's sequence handed out so far
    bool more = base < n;
    while (true) {
        const uint64_t idle = __ballot(!live);
        if (more && (idle == ~0ull || uint32_t(__popcll(idle)) >= kRefill)) {
            if (!live) {
                const uint32_t pos = next + uint32_t(__popcll(idle & lanemask_lt()));
                const uint64_t slot = uint64_t(pos / kWave) * stride + base + pos % kWave;
                if (slot < n) {
                    qi = uint32_t(slot);
                    i = queue[qi];
                    if (MULTI) T = tree_of(jw[ray_job[i >> job_shift]]);
                    live = wide_ray_start(R, rays, i);
                    if (!live) wide_ray_finish(R, i, prim_out, t_out, bary_out);
                }
            }
            next += uint32_t(__popcll(idle));
            more = uint64_t(next / kWave) * stride + base + next % kWave < n;
            continue;
        }
        if (idle == ~0ull) break;  // nothing in flight and nothing left
        if (live) {
            live = wide_step<COUNT, SHAPES>(R, T, B, any_hit, qi, err, s_stack, visits, boxes, tests);
            if (!live) wide_ray_finish(R, i, prim_out, t_out, bary_out);
        }
    }
    add_counts<COUNT>(counts, visits, boxes, tests);
}

// The occupancy targets and the refill threshold come from measurement: the
// persistent single-tree traversal runs fastest at 5 waves per SIMD with refills
// at 8 idle lanes (6 waves: -10 %; 16 lanes: -4 %), the per-ray-tree one at 5;
// the LDS stack (24 KiB per block) allows 6.
// SHAPES: trees with analytic primitives (a leaf position whose pad word holds a
// shape index + 1); triangle-only trees take the instance without that branch.
template <bool COUNT, bool SHAPES>
__global__ __launch_bounds__(kWideBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void trace_wide_kernel(
    WideTrace B, const float4 *__restrict__ rays, const uint32_t *__restrict__ queue,
    const uint32_t *__restrict__ queue_len, int any_hit, uint32_t *__restrict__ prim_out,
    float *__restrict__ t_out, float *__restrict__ bary_out, uint32_t *__restrict__ err,
    unsigned long long *__restrict__ counts) {
    __shared__ uint2 s_stack[kLdsStack][kWideBlock];
    trace_wide_persistent_body<COUNT, false, SHAPES>(B, rays, queue, queue_len, any_hit, prim_out, t_out, bary_out,
                                                     err, counts, nullptr, nullptr, 0, s_stack);
}

template <bool SHAPES>
__global__ __launch_bounds__(kWideBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void trace_wide_multi_kernel(
    WideTrace B, const float4 *__restrict__ rays, const uint32_t *__restrict__ queue,
    const uint32_t *__restrict__ queue_len, int any_hit, uint32_t *__restrict__ prim_out,
    float *__restrict__ t_out, float *__restrict__ bary_out, uint32_t *__restrict__ err,
    const WideTrace *__restrict__ jw, const uint32_t *__restrict__ ray_job, uint32_t job_shift) {
    __shared__ uint2 s_stack[kLdsStack][kWideBlock];
    trace_wide_persistent_body<false, true, SHAPES>(B, rays, queue, queue_len, any_hit, prim_out, t_out, bary_out,
                                                    err, nullptr, jw, ray_job, job_shift, s_stack);
}

}  // namespace

hipError_t collapse_wide(const BNode *bnodes, uint32_t n, WideBvh &out, DevBuf &scratch, uint32_t *err,
                         hipStream_t stream) {
    out.num_nodes = out.max_depth = out.collapse_levels = 0;
    if (n == 0) return hipSuccess;
    const uint32_t cap = n;  // one wide node per binary interior node (n - 1), or one for a leaf root
    hipError_t e = out.nodes.reserve(size_t(cap) * sizeof(WNode));
    if (e == hipSuccess) e = out.order.reserve(size_t(cap) * 8 * sizeof(uint32_t));
    const size_t job_bytes = size_t(cap) * sizeof(Job);
    if (e == hipSuccess) e = scratch.reserve(2 * job_bytes + 256);
    if (e != hipSuccess) return e;
    Job *jobs[2] = {reinterpret_cast<Job *>(scratch.ptr),
                    reinterpret_cast<Job *>(static_cast<char *>(scratch.ptr) + job_bytes)};
    uint32_t *counters = reinterpret_cast<uint32_t *>(static_cast<char *>(scratch.ptr) + 2 * job_bytes);
    // counters: [4] levels that had jobs, [kCollapseJobs + L % 3] the job count of
    // level L, [kCollapseBase + L % 3] the wide index of its first job
    e = hipMemsetAsync(counters, 0, 64, stream);
    if (e != hipSuccess) return e;
    WNode *wn = out.nodes.as<WNode>();
    uint32_t *order = out.order.as<uint32_t>();
    wide_root_kernel<<<1, 64, 0, stream>>>(bnodes, n, jobs[0], counters, wn);
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wide_collapse_kernel, kCollapseBlock, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu <= 0 ||
        cus <= 0) {
        per_cu = 4;
        cus = 256;
    }
    const uint32_t resident = uint32_t(per_cu) * uint32_t(cus);
    // levels are enqueued in batches between reads of the job count (the first
    // batch: about the depth of a full 8-wide tree over n leaves, plus two); a
    // level past the last one finds no jobs and exits at once
    uint32_t host[kCollapseBase + 3];
    uint32_t level = 0;
    int cur = 0;
    int batch = 2;
    for (uint64_t c = 1; c < n; c <<= 3) ++batch;
    for (;;) {
        for (int k = 0; k < batch; ++k, ++level) {
            if (level > 100000) return hipErrorUnknown;
            wide_collapse_kernel<<<resident, kCollapseBlock, 0, stream>>>(bnodes, jobs[cur], jobs[1 - cur], counters,
                                                                          wn, order, cap, 2 * n - 1, err, level);
            cur = 1 - cur;
        }
        batch = 4;
        e = hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return e;
        if (host[kCollapseJobs + level % 3] == 0) break;  // the next level has no jobs
    }
    out.collapse_levels = host[4];
    out.num_nodes = host[kCollapseBase + level % 3];  // B of the first level without jobs
    // every job of a level has the same depth, so the depth of the wide tree is
    // the number of levels that had jobs (a leaf root: one level, no jobs)
    out.max_depth = std::max(host[4], 1u);
    return hipGetLastError();
}

// Workgroups of a persistent traversal kernel: as many as are resident on the
// device at once (quarters/4 of them when quarters is 1..3), at most one per 256
// rays. The resident count is computed once per (kernel, device).
template <typename Kernel>
static uint32_t persistent_grid(Kernel kernel, uint32_t n, uint32_t quarters) {
    struct Entry {
        const void *kernel;
        int device;
        uint32_t blocks;
    };
    static thread_local Entry cache[8] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const void *key = reinterpret_cast<const void *>(kernel);
    uint32_t blocks = 0;
    for (const Entry &e : cache)
        if (e.kernel == key && e.device == dev) blocks = e.blocks;
    if (blocks == 0) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWideBlock, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu <= 0 ||
            cus <= 0) {
            per_cu = 1;
            cus = 256;
        }
        blocks = uint32_t(per_cu) * uint32_t(cus);
        for (Entry &e : cache)
            if (!e.kernel) {
                e = Entry{key, dev, blocks};
                break;
            }
    }
    if (quarters >= 1 && quarters < 4) blocks = std::max(1u, blocks * quarters / 4);
    return std::max(1u, std::min(blocks, (n + kWideBlock - 1) / kWideBlock));
}

// The persistent traversal over a compacted queue; the instance with the
// analytic-primitive branch when the tree has analytic primitives.
static hipError_t trace_wide_queued_counts(const WideTrace &B, const float4 *r, uint32_t n, int any_hit,
                                           uint32_t *prim_out, float *t_out, float *bary_out, uint32_t *err,
                                           unsigned long long *counts, const uint32_t *queue,
                                           const uint32_t *queue_len, hipStream_t stream) {
    const uint32_t grid = persistent_grid(trace_wide_kernel<false, false>, n, B.grid_quarters);
    if (B.shapes) {
        if (counts)
            trace_wide_kernel<true, true><<<grid, kWideBlock, 0, stream>>>(B, r, queue, queue_len, any_hit, prim_out,
                                                                           t_out, bary_out, err, counts);
        else
            trace_wide_kernel<false, true><<<grid, kWideBlock, 0, stream>>>(B, r, queue, queue_len, any_hit,
                                                                            prim_out, t_out, bary_out, err, nullptr);
    } else if (counts) {
        trace_wide_kernel<true, false><<<grid, kWideBlock, 0, stream>>>(B, r, queue, queue_len, any_hit, prim_out,
                                                                        t_out, bary_out, err, counts);
    } else {
        trace_wide_kernel<false, false><<<grid, kWideBlock, 0, stream>>>(B, r, queue, queue_len, any_hit, prim_out,
                                                                         t_out, bary_out, err, nullptr);
    }
    return hipGetLastError();
}

// B.shapes non-null: some job's tree has analytic primitives.
static hipError_t launch_multi(const WideTrace &B, const WideTrace *jobs, const uint32_t *ray_job,
                               uint32_t job_shift, const float4 *r, uint32_t n, int any_hit, uint32_t *prim_out,
                               float *t_out, float *bary_out, uint32_t *err, const uint32_t *queue,
                               const uint32_t *queue_len, hipStream_t stream) {
    const uint32_t grid = persistent_grid(trace_wide_multi_kernel<false>, n, B.grid_quarters);
    if (B.shapes)
        trace_wide_multi_kernel<true><<<grid, kWideBlock, 0, stream>>>(B, r, queue, queue_len, any_hit, prim_out,
                                                                       t_out, bary_out, err, jobs, ray_job,
                                                                       job_shift);
    else
        trace_wide_multi_kernel<false><<<grid, kWideBlock, 0, stream>>>(B, r, queue, queue_len, any_hit, prim_out,
                                                                        t_out, bary_out, err, jobs, ray_job,
                                                                        job_shift);
    return hipGetLastError();
}

hipError_t trace_wide(const WideTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                      float *t_out, float *bary_out, uint32_t *err, unsigned long long *counts, uint32_t *queue,
                      uint32_t *queue_len, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t cgrid = (n + kCompactBlock * kCompactItems - 1) / (kCompactBlock * kCompactItems);
    const float4 *r = reinterpret_cast<const float4 *>(rays);
    compact_rays_kernel<<<cgrid, kCompactBlock, 0, stream>>>(r, n, queue, queue_len, prim_out, t_out, bary_out);
    return trace_wide_queued_counts(B, r, n, any_hit, prim_out, t_out, bary_out, err, counts, queue, queue_len,
                                    stream);
}

namespace {

__device__ __forceinline__ uint32_t spread3_6(uint32_t v) {  // 6 bits -> every third bit
    v &= 0x3Fu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// MULTI: rays of a batch; the ray's job (low 10 bits) leads the key so that a
// wave's rays share a tree, and the keys take three radix passes.
template <bool MULTI>
__global__ __launch_bounds__(256) void ray_keys_kernel(const float4 *__restrict__ rays, uint32_t n, WideTrace B,
                                                       uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                       uint32_t *__restrict__ hist, uint32_t *__restrict__ count,
                                                       uint32_t *__restrict__ prim_out, float *__restrict__ t_out,
                                                       float *__restrict__ bary_out,
                                                       const uint32_t *__restrict__ ray_job, uint32_t job_shift) {
    __shared__ uint32_t s_hist[kSortPasses * kSortRadix];
    for (int i = threadIdx.x; i < int(kSortPasses * kSortRadix); i += blockDim.x) s_hist[i] = 0u;
    __syncthreads();
    const float sx = B.hi[0] > B.lo[0] ? 64.f / (B.hi[0] - B.lo[0]) : 0.f;
    const float sy = B.hi[1] > B.lo[1] ? 64.f / (B.hi[1] - B.lo[1]) : 0.f;
    const float sz = B.hi[2] > B.lo[2] ? 64.f / (B.hi[2] - B.lo[2]) : 0.f;
    uint32_t mine = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 r0 = rays[2 * size_t(i)];
        // inactive: after every active ray (active keys are < 2^19, or < 2^30 with jobs)
        uint32_t key = MULTI ? 0x3FFFFFFFu : 0xFFFFFu;
        if (r0.w > 0.f) {
            const float4 r1 = rays[2 * size_t(i) + 1];
            const uint32_t qx = uint32_t(fminf(63.f, fmaxf(0.f, (r0.x - B.lo[0]) * sx)));
            const uint32_t qy = uint32_t(fminf(63.f, fmaxf(0.f, (r0.y - B.lo[1]) * sy)));
            const uint32_t qz = uint32_t(fminf(63.f, fmaxf(0.f, (r0.z - B.lo[2]) * sz)));
            const uint32_t oct = (r1.x < 0.f ? 1u : 0u) | (r1.y < 0.f ? 2u : 0u) | (r1.z < 0.f ? 4u : 0u);
            const uint32_t m = (spread3_6(qx) << 2) | (spread3_6(qy) << 1) | spread3_6(qz);  // 18 bits
            key = (oct << 16) | (m >> 2);  // < 2^19: below the inactive key, two sort passes
            if (MULTI) key |= (ray_job[i >> job_shift] & 1023u) << 19;
            ++mine;
        } else {
            prim_out[i] = kNoHit;
            if (t_out) t_out[i] = 0.f;
            if (bary_out) {
                bary_out[3 * size_t(i)] = 0.f;
                bary_out[3 * size_t(i) + 1] = 0.f;
                bary_out[3 * size_t(i) + 2] = 0.f;
            }
        }
        keys[i] = key;
        vals[i] = i;
        atomicAdd(&s_hist[key & 1023u], 1u);
        atomicAdd(&s_hist[kSortRadix + ((key >> 10) & 1023u)], 1u);
        if (MULTI) atomicAdd(&s_hist[2 * kSortRadix + ((key >> 20) & 1023u)], 1u);
    }
    const uint32_t tot = wave_sum(mine);
    if (__lane_id() == 0 && tot) atomicAdd(count, tot);
    __syncthreads();
    for (int i = threadIdx.x; i < int(kSortPasses * kSortRadix); i += blockDim.x)
        if (s_hist[i]) atomicAdd(&hist[i], s_hist[i]);
}

__global__ void zero_hist_kernel(uint32_t *h) {
    for (uint32_t i = threadIdx.x; i < kSortPasses * kSortRadix; i += blockDim.x) h[i] = 0u;
}

}  // namespace

hipError_t reserve_ray_sort(RaySortScratch &s, uint32_t n) {
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = s.keys[k].reserve(size_t(n) * 4);
        if (e == hipSuccess) e = s.vals[k].reserve(size_t(n) * 4);
    }
    if (e == hipSuccess) e = s.hist.reserve(kSortPasses * kSortRadix * 4);
    if (e == hipSuccess) e = onesweep_reserve(s.sort, n);
    return e;
}

hipError_t sort_rays(const WideTrace &B, const float *rays, uint32_t n, uint32_t *prim_out, float *t_out,
                     float *bary_out, uint32_t *err, RaySortScratch &s, uint32_t *queue_len, uint32_t **queue,
                     hipStream_t stream, const uint32_t *ray_job, uint32_t job_shift) {
    hipError_t e = reserve_ray_sort(s, n);
    if (e != hipSuccess) return e;
    *queue = s.vals[0].as<uint32_t>();
    if (n == 0) return hipSuccess;
    zero_hist_kernel<<<1, 256, 0, stream>>>(s.hist.as<uint32_t>());  // a kernel, so that a graph may capture it
    const uint32_t grid = std::min((n + 255) / 256, 1024u);
    const float4 *r = reinterpret_cast<const float4 *>(rays);
    if (ray_job)
        ray_keys_kernel<true><<<grid, 256, 0, stream>>>(r, n, B, s.keys[0].as<uint32_t>(), s.vals[0].as<uint32_t>(),
                                                        s.hist.as<uint32_t>(), queue_len, prim_out, t_out, bary_out,
                                                        ray_job, job_shift);
    else
        ray_keys_kernel<false><<<grid, 256, 0, stream>>>(r, n, B, s.keys[0].as<uint32_t>(), s.vals[0].as<uint32_t>(),
                                                         s.hist.as<uint32_t>(), queue_len, prim_out, t_out,
                                                         bary_out, nullptr, 0);
    uint32_t *kr = nullptr, *vr = nullptr;
    e = onesweep_sort(s.keys[0].as<uint32_t>(), s.vals[0].as<uint32_t>(), s.keys[1].as<uint32_t>(),
                      s.vals[1].as<uint32_t>(), n, s.hist.as<uint32_t>(), err, s.sort, stream, &kr, &vr,
                      ray_job ? 3u : 2u);
    *queue = vr;
    return e;
}

hipError_t trace_wide_multi_queued(const WideTrace &B, const WideTrace *jobs, const uint32_t *ray_job,
                                   uint32_t job_shift, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                                   float *t_out, float *bary_out, uint32_t *err, const uint32_t *queue,
                                   const uint32_t *queue_len, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    return launch_multi(B, jobs, ray_job, job_shift, reinterpret_cast<const float4 *>(rays), n, any_hit, prim_out,
                        t_out, bary_out, err, queue, queue_len, stream);
}

hipError_t trace_wide_queued(const WideTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                             float *t_out, float *bary_out, uint32_t *err, const uint32_t *queue,
                             const uint32_t *queue_len, hipStream_t stream, unsigned long long *counts) {
    if (n == 0) return hipSuccess;
    return trace_wide_queued_counts(B, reinterpret_cast<const float4 *>(rays), n, any_hit, prim_out, t_out, bary_out,
                                    err, counts, queue, queue_len, stream);
}

hipError_t trace_wide_multi(const WideTrace &B, const WideTrace *jobs, const uint32_t *ray_job, uint32_t job_shift,
                            const float *rays, uint32_t n, int any_hit, uint32_t *prim_out, float *t_out,
                            float *bary_out, uint32_t *err, uint32_t *queue, uint32_t *queue_len,
                            hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t cgrid = (n + kCompactBlock * kCompactItems - 1) / (kCompactBlock * kCompactItems);
    const float4 *r = reinterpret_cast<const float4 *>(rays);
    compact_rays_kernel<<<cgrid, kCompactBlock, 0, stream>>>(r, n, queue, queue_len, prim_out, t_out, bary_out);
    return launch_multi(B, jobs, ray_job, job_shift, r, n, any_hit, prim_out, t_out, bary_out, err, queue, queue_len,
                        stream);
}

}  // namespace pbrtm
