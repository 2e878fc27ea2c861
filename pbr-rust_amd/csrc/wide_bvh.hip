// 8-wide BVH: collapse of the binary build tree and the closest/any-hit
// traversal over it.
//
// Collapse: each wide node stands for one binary interior node; its frontier
// is grown by repeatedly opening the frontier interior node of largest surface
// area until 8 slots are filled (the greedy collapse of Wald et al. 2008 /
// Ylitie et al. 2017). Child boxes are the binary nodes' exact FP32 boxes.
//
// Traversal keeps the reference's semantics exactly (accelerator/hlbvh.cu:179-289):
// for every ray octant a wide node stores the order in which the binary
// depth-first, near-child-first traversal reaches its slots, and a slot is
// entered only if its box passes the slab test against the current best t at
// the moment the binary traversal would pop it. Box tests are monotone under
// box inclusion, so skipping the collapsed inner binary boxes changes no
// decision: the sequence of triangle tests, and therefore every hit (ties
// included), equals the binary traversal's. Each visit costs one 256-byte node
// (two cache lines) instead of up to seven dependent 32-byte fetches.
#include "pbrtm_geometry.h"
#include "wave_ops.h"
#include "wide_bvh.h"

namespace pbrtm {

namespace {

struct Job {
    uint32_t bnode, wnode, depth, pad;
};

__device__ __forceinline__ Box bnode_box(const BNode &b) {
    return {{b.lo[0], b.lo[1], b.lo[2]}, {b.hi[0], b.hi[1], b.hi[2]}};
}
__device__ __forceinline__ bool bnode_leaf(const BNode &b) { return (b.a & kLeafFlag) != 0; }

__global__ void wide_root_kernel(const BNode *bn, Job *jobs, uint32_t *counters, WNode *wn) {
    if (threadIdx.x != 0) return;
    const BNode r = bn[0];
    if (bnode_leaf(r)) {  // the whole tree is one leaf: a wide root with one leaf slot
        WNode w;
        for (int s = 0; s < kWide; ++s) {
            const bool use = s == 0;
            w.lo_x[s] = use ? r.lo[0] : 3.402823466e+38f; w.lo_y[s] = use ? r.lo[1] : 3.402823466e+38f;
            w.lo_z[s] = use ? r.lo[2] : 3.402823466e+38f;
            w.hi_x[s] = use ? r.hi[0] : -3.402823466e+38f; w.hi_y[s] = use ? r.hi[1] : -3.402823466e+38f;
            w.hi_z[s] = use ? r.hi[2] : -3.402823466e+38f;
            w.ref[s][0] = use ? r.a : kWideEmpty;
            w.ref[s][1] = use ? r.b : 0u;
        }
        wn[0] = w;
        counters[1] = 1;  // nodes used
        counters[0] = 0;  // no jobs
        return;
    }
    jobs[0] = Job{0, 0, 1, 0};
    counters[0] = 1;
    counters[1] = 1;
}

__global__ __launch_bounds__(64) void wide_collapse_kernel(const BNode *__restrict__ bn, const Job *__restrict__ jobs,
                                                           uint32_t njobs, Job *__restrict__ next,
                                                           uint32_t *__restrict__ counters, WNode *__restrict__ wn,
                                                           uint32_t *__restrict__ order, uint32_t cap,
                                                           uint32_t *__restrict__ err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < njobs;
    Job job = live ? jobs[i] : Job{0, 0, 0, 0};
    uint32_t F[kWide];   // frontier binary ids, left to right
    uint32_t E[kWide];   // opened binary interior nodes (the collapsed subtree's inner nodes)
    int nF = 0, nE = 0;
    if (live) {
        const BNode r = bn[job.bnode];
        E[nE++] = job.bnode;
        F[nF++] = r.a;
        F[nF++] = r.b & 0x3FFFFFFFu;
        while (nF < kWide) {
            int best = -1;
            float best_area = -1.0f;
            for (int k = 0; k < nF; ++k) {
                const BNode c = bn[F[k]];
                if (bnode_leaf(c)) continue;
                const float a = box_area(bnode_box(c));
                if (a > best_area) { best_area = a; best = k; }
            }
            if (best < 0) break;
            const BNode c = bn[F[best]];
            E[nE++] = F[best];
            for (int k = nF; k > best + 1; --k) F[k] = F[k - 1];
            F[best] = c.a;
            F[best + 1] = c.b & 0x3FFFFFFFu;
            ++nF;
        }
    }
    // interior children need wide node indices: one atomic per wave
    uint32_t n_int = 0;
    bool is_int[kWide];
    for (int s = 0; s < kWide; ++s) {
        is_int[s] = live && s < nF && !bnode_leaf(bn[F[s]]);
        n_int += is_int[s] ? 1u : 0u;
    }
    const uint32_t incl = wave_inclusive_add(n_int);
    uint32_t base = 0;
    if (__lane_id() == kWave - 1 && incl > 0) base = atomicAdd(&counters[1], incl);
    base = __shfl(base, kWave - 1, kWave) + incl - n_int;
    uint32_t jbase = 0;
    if (__lane_id() == kWave - 1 && incl > 0) jbase = atomicAdd(&counters[2], incl);
    jbase = __shfl(jbase, kWave - 1, kWave) + incl - n_int;
    if (!live) return;
    if (base + n_int > cap) {
        atomicOr(err, uint32_t(kErrCapacity));
        return;
    }
    WNode w;
    uint32_t k_int = 0;
    for (int s = 0; s < kWide; ++s) {
        if (s < nF) {
            const BNode c = bn[F[s]];
            w.lo_x[s] = c.lo[0]; w.lo_y[s] = c.lo[1]; w.lo_z[s] = c.lo[2];
            w.hi_x[s] = c.hi[0]; w.hi_y[s] = c.hi[1]; w.hi_z[s] = c.hi[2];
            if (is_int[s]) {
                const uint32_t idx = base + k_int;
                next[jbase + k_int] = Job{F[s], idx, job.depth + 1, 0};
                ++k_int;
                w.ref[s][0] = idx;
                w.ref[s][1] = 0u;
            } else {
                w.ref[s][0] = c.a;  // kLeafFlag | first (same bit as kWideLeaf)
                w.ref[s][1] = c.b;
            }
        } else {
            w.lo_x[s] = w.lo_y[s] = w.lo_z[s] = 3.402823466e+38f;
            w.hi_x[s] = w.hi_y[s] = w.hi_z[s] = -3.402823466e+38f;
            w.ref[s][0] = kWideEmpty;
            w.ref[s][1] = 0u;
        }
    }
    wn[job.wnode] = w;
    // depth-first, near-first slot order per octant (the binary traversal's order)
    for (uint32_t oct = 0; oct < 8; ++oct) {
        uint32_t st[kWide];
        int sp = 0;
        st[sp++] = job.bnode;
        uint32_t perm = 0xFFFFFFFFu;
        int pos = 0;
        while (sp > 0) {
            const uint32_t x = st[--sp];
            bool opened = false;
            for (int k = 0; k < nE; ++k) opened |= E[k] == x;
            if (opened) {
                const BNode b = bn[x];
                const uint32_t axis = b.b >> 30;
                const bool neg = (oct >> axis) & 1u;
                const uint32_t l = b.a, r = b.b & 0x3FFFFFFFu;
                st[sp++] = neg ? l : r;  // far child waits
                st[sp++] = neg ? r : l;
            } else {
                uint32_t slot = 0;
                for (int k = 0; k < nF; ++k)
                    if (F[k] == x) slot = k;
                perm = (perm & ~(0xFu << (4 * pos))) | (slot << (4 * pos));
                ++pos;
            }
        }
        order[job.wnode * 8 + oct] = perm;
    }
    atomicMax(&counters[3], job.depth);
}

// Entry distance of one child box, or +inf when the slab test fails for a
// reason that does not depend on the ray's current t_max (bounds3.h:146-195).
// box_hit(..., t) == (entry < t).
__device__ __forceinline__ float box_entry(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv,
                                           const int neg[3]) {
    const double widen = 1.0 + 2.0 * double(gamma_n(3));
    float t_min = ((neg[0] ? hx : lx) - o.x) * inv.x;
    float t_max = ((neg[0] ? lx : hx) - o.x) * inv.x;
    const float ty_min = ((neg[1] ? hy : ly) - o.y) * inv.y;
    float ty_max = ((neg[1] ? ly : hy) - o.y) * inv.y;
    t_max = float(double(t_max) * widen);
    ty_max = float(double(ty_max) * widen);
    if (t_min > ty_max || ty_min > t_max) return __builtin_inff();
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    const float tz_min = ((neg[2] ? hz : lz) - o.z) * inv.z;
    float tz_max = ((neg[2] ? lz : hz) - o.z) * inv.z;
    tz_max = float(double(tz_max) * widen);
    if (t_min > tz_max || tz_min > t_max) return __builtin_inff();
    if (tz_min > t_min) t_min = tz_min;
    if (tz_max < t_max) t_max = tz_max;
    return t_max > 0.0f ? t_min : __builtin_inff();
}

constexpr int kWideBlock = 256;

template <bool COUNT>
__global__ __launch_bounds__(kWideBlock) void trace_wide_kernel(WideTrace B, const float4 *__restrict__ rays,
                                                                uint32_t n, int any_hit,
                                                                uint32_t *__restrict__ prim_out,
                                                                float *__restrict__ t_out,
                                                                float *__restrict__ bary_out,
                                                                uint32_t *__restrict__ err,
                                                                unsigned long long *__restrict__ counts) {
    __shared__ uint2 s_stack[kLdsStack][kWideBlock];
    const uint32_t i = blockIdx.x * kWideBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    const f3 o = {r0.x, r0.y, r0.z}, d = {r1.x, r1.y, r1.z};
    float best_t = r0.w;
    uint32_t best = kNoHit;
    TriHit hit{0.f, 0.f, 0.f, 0.f};
    uint32_t visits = 0, boxes = 0, tests = 0;
    if (best_t > 0.f) {
        const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        const int neg[3] = {inv.x < 0.f, inv.y < 0.f, inv.z < 0.f};
        const uint32_t oct = uint32_t(neg[0]) | (uint32_t(neg[1]) << 1) | (uint32_t(neg[2]) << 2);
        uint32_t sp = 0;
        uint32_t node = 0, pmask = 0xFFu;
        bool done = false;
        while (!done) {
            if (node >= B.num_nodes) {  // only a corrupt layout gets here
                atomicOr(err, uint32_t(kErrStack));
                break;
            }
            const WNode *W = B.nodes + node;
            const uint32_t perm = B.order[node * 8 + oct];
            if (COUNT) { ++visits; boxes += kWide; }
            // all eight child boxes (12 x 16-byte loads), hit mask in slot order
            const float4 *q = reinterpret_cast<const float4 *>(W);
            float pl[6][kWide];
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                const float4 v = q[k];
                pl[k / 2][(k % 2) * 4 + 0] = v.x;
                pl[k / 2][(k % 2) * 4 + 1] = v.y;
                pl[k / 2][(k % 2) * 4 + 2] = v.z;
                pl[k / 2][(k % 2) * 4 + 3] = v.w;
            }
            uint32_t hs = 0;
#pragma unroll
            for (int s = 0; s < kWide; ++s) {
                const float e = box_entry(pl[0][s], pl[1][s], pl[2][s], pl[3][s], pl[4][s], pl[5][s], o, inv, neg);
                hs |= (e < best_t) ? (1u << s) : 0u;
            }
            // ... then in depth-first position order
            uint32_t hm = 0;
#pragma unroll
            for (int p = 0; p < kWide; ++p) {
                const uint32_t s = (perm >> (4 * p)) & 0xFu;
                hm |= (s < kWide && ((hs >> s) & 1u)) ? (1u << p) : 0u;
            }
            pmask &= hm;
            const float t_at_visit = best_t;
            bool descended = false;
            while (pmask) {
                const uint32_t p = __builtin_ctz(pmask);
                pmask &= pmask - 1;
                const uint32_t s = (perm >> (4 * p)) & 0xFu;
                if (best_t != t_at_visit) {  // a leaf of this node shortened the ray: re-test at pop time
                    if (COUNT) ++boxes;
                    const float e = box_entry(W->lo_x[s], W->lo_y[s], W->lo_z[s], W->hi_x[s], W->hi_y[s],
                                              W->hi_z[s], o, inv, neg);
                    if (!(e < best_t)) continue;
                }
                const uint32_t ref0 = W->ref[s][0];
                if (ref0 & kWideLeaf) {
                    const uint32_t first = ref0 & ~kWideLeaf, cnt = W->ref[s][1];
                    if (uint64_t(first) + cnt > B.num_prims) {
                        atomicOr(err, uint32_t(kErrStack));
                        done = true;
                        break;
                    }
                    for (uint32_t k = first; k < first + cnt; ++k) {
                        const float *t = B.tri_leaf + 9ull * k;
                        if (COUNT) ++tests;
                        TriHit h;
                        if (tri_hit(o, d, best_t, f3{t[0], t[1], t[2]}, f3{t[3], t[4], t[5]}, f3{t[6], t[7], t[8]},
                                    &h)) {
                            best = B.prim_order[k];
                            hit = h;
                            best_t = h.t;
                            if (any_hit) { done = true; break; }
                        }
                    }
                    if (done) break;
                    continue;
                }
                if (pmask) {  // the rest of this node waits
                    if (sp < kLdsStack) {
                        s_stack[sp][threadIdx.x] = make_uint2(node, pmask);
                    } else if (sp - kLdsStack < B.spill_per_ray) {
                        B.spill[uint64_t(i) * B.spill_per_ray + (sp - kLdsStack)] = make_uint2(node, pmask);
                    } else {
                        atomicOr(err, uint32_t(kErrStack));
                        done = true;
                        break;
                    }
                    ++sp;
                }
                node = ref0;
                pmask = 0xFFu;
                descended = true;
                break;
            }
            if (done || descended) continue;
            if (sp == 0) break;
            --sp;
            const uint2 e = sp < kLdsStack ? s_stack[sp][threadIdx.x]
                                           : B.spill[uint64_t(i) * B.spill_per_ray + (sp - kLdsStack)];
            node = e.x;
            pmask = e.y;
        }
    }
    if (COUNT) {
        const unsigned long long v = wave_sum(visits), b = wave_sum(boxes), tt = wave_sum(tests);
        if (__lane_id() == 0) {
            atomicAdd(&counts[0], v);
            atomicAdd(&counts[1], b);
            atomicAdd(&counts[2], tt);
        }
    }
    prim_out[i] = best;
    if (t_out) t_out[i] = best == kNoHit ? 0.f : hit.t;
    if (bary_out) {
        bary_out[3 * i] = hit.b0;
        bary_out[3 * i + 1] = hit.b1;
        bary_out[3 * i + 2] = hit.b2;
    }
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
    // counters: [0] jobs this level, [1] wide nodes used, [2] jobs next level, [3] max depth
    e = hipMemsetAsync(counters, 0, 64, stream);
    if (e != hipSuccess) return e;
    WNode *wn = out.nodes.as<WNode>();
    uint32_t *order = out.order.as<uint32_t>();
    wide_root_kernel<<<1, 64, 0, stream>>>(bnodes, jobs[0], counters, wn);
    uint32_t host[4];
    int cur = 0;
    for (;;) {
        e = hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return e;
        const uint32_t nj = host[0];
        if (nj == 0) break;
        if (++out.collapse_levels > 100000) return hipErrorUnknown;
        e = hipMemsetAsync(counters + 2, 0, 4, stream);
        if (e != hipSuccess) return e;
        wide_collapse_kernel<<<(nj + 63) / 64, 64, 0, stream>>>(bnodes, jobs[cur], nj, jobs[1 - cur], counters, wn,
                                                                 order, cap, err);
        e = hipMemcpyAsync(counters, counters + 2, 4, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return e;
        cur = 1 - cur;
    }
    out.num_nodes = host[1];
    out.max_depth = std::max(host[3], 1u);
    return hipGetLastError();
}

hipError_t trace_wide(const WideTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                      float *t_out, float *bary_out, uint32_t *err, unsigned long long *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (n + kWideBlock - 1) / kWideBlock;
    const float4 *r = reinterpret_cast<const float4 *>(rays);
    if (counts)
        trace_wide_kernel<true><<<grid, kWideBlock, 0, stream>>>(B, r, n, any_hit, prim_out, t_out, bary_out, err,
                                                                 counts);
    else
        trace_wide_kernel<false><<<grid, kWideBlock, 0, stream>>>(B, r, n, any_hit, prim_out, t_out, bary_out, err,
                                                                  nullptr);
    return hipGetLastError();
}

}  // namespace pbrtm
