// Closest-hit and any-hit traversal (accelerator/hlbvh.cu:179-289, reference)
// over the binary node array, plus the ambient-occlusion bounce used by the
// primary + 1 diffuse bounce workload (integrators/ambient_occlusion.cu:28-52).
//
// The binary traversal visits nodes in the reference's order: depth first,
// near child first by the sign of the ray direction on the node's split axis,
// with the same slab test. Leaves read triangles stored contiguously in leaf
// order (36 B each), so a leaf costs one coalesced span instead of the
// reference's primitive -> shape -> mesh pointer chase.
#include "traverse.h"
#include "pbrtm_geometry.h"
#include "wave_ops.h"

namespace pbrtm {

namespace {

constexpr int kTraceBlock = 64;
constexpr int kStack = 64;

__device__ __forceinline__ f3 load3(const float *p) { return {p[0], p[1], p[2]}; }

// COUNT adds per-launch totals of box tests and triangle tests (for the
// algorithmic-bytes roofline); the traversal itself is identical.
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void trace_binary_kernel(
    const BNode *__restrict__ nodes, uint32_t num_nodes, uint32_t num_prims, const float *__restrict__ tri_leaf,
    const uint32_t *__restrict__ prim_order, const float4 *__restrict__ rays, uint32_t n, int any_hit,
    uint32_t *__restrict__ prim_out, float *__restrict__ t_out, float *__restrict__ bary_out,
    uint32_t *__restrict__ err, unsigned long long *__restrict__ counts) {
    __shared__ uint32_t s_stack[kStack][kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    const f3 o = {r0.x, r0.y, r0.z}, d = {r1.x, r1.y, r1.z};
    float best_t = r0.w;
    uint32_t best = kNoHit;
    TriHit hit{0.f, 0.f, 0.f, 0.f};
    uint32_t visits = 0, tests = 0;
    if (best_t > 0.f) {
        const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        const int neg[3] = {inv.x < 0.f, inv.y < 0.f, inv.z < 0.f};
        int sp = 0;
        uint32_t cur = 0;
        for (;;) {
            if (cur >= num_nodes) {  // only a corrupt tree gets here
                atomicOr(err, uint32_t(kErrStack));
                break;
            }
            const BNode nd = nodes[cur];
            if (COUNT) ++visits;
            if (box_hit(nd.lo, nd.hi, o, inv, neg, best_t)) {
                if (nd.a & kLeafFlag) {
                    const uint32_t first = nd.a & ~kLeafFlag;
                    if (uint64_t(first) + nd.b > num_prims) {
                        atomicOr(err, uint32_t(kErrStack));
                        break;
                    }
                    bool done = false;
                    for (uint32_t k = first; k < first + nd.b; ++k) {
                        const float4 *t4 = reinterpret_cast<const float4 *>(tri_leaf) + 3ull * k;
                        const float4 q0 = t4[0], q1 = t4[1], q2 = t4[2];
                        if (COUNT) ++tests;
                        TriHit h;
                        if (tri_hit(o, d, best_t, f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, f3{q1.z, q1.w, q2.x},
                                    &h)) {
                            best = prim_order[k];
                            hit = h;
                            best_t = h.t;
                            if (any_hit) { done = true; break; }
                        }
                    }
                    if (done) break;
                } else {
                    const uint32_t axis = nd.b >> 30;
                    const uint32_t left = nd.a, right = nd.b & 0x3FFFFFFFu;
                    const bool rev = neg[axis];
                    if (sp >= kStack) {
                        atomicOr(err, uint32_t(kErrStack));
                        break;
                    }
                    s_stack[sp++][threadIdx.x] = rev ? left : right;  // far child waits
                    cur = rev ? right : left;
                    continue;
                }
            }
            if (sp == 0) break;
            cur = s_stack[--sp][threadIdx.x];
        }
    }
    if (COUNT) {
        const unsigned long long v = wave_sum(visits), tt = wave_sum(tests);
        if (__lane_id() == 0) {
            atomicAdd(&counts[0], v);
            atomicAdd(&counts[1], v);
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

__global__ __launch_bounds__(256) void deindex_kernel(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                                      uint32_t nt, uint32_t nv, float *__restrict__ tri,
                                                      uint32_t *__restrict__ err) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        uint32_t j = idx[3 * t + v];
        if (j >= nv) {
            atomicOr(err, uint32_t(kErrNonFinite));
            j = 0;
        }
        tri[9ull * t + 3 * v] = verts[3ull * j];
        tri[9ull * t + 3 * v + 1] = verts[3ull * j + 1];
        tri[9ull * t + 3 * v + 2] = verts[3ull * j + 2];
    }
}

// Cosine-hemisphere bounce from each closest hit (ambient_occlusion.cu:36-52):
// u = two PCG32 draws of sequence = ray index, around the face-forwarded
// geometric normal; the origin is offset past the hit point's error box.
__global__ __launch_bounds__(256) void ao_bounce_kernel(const float *__restrict__ tri, const float4 *__restrict__ rays,
                                                        const uint32_t *__restrict__ prim,
                                                        const float *__restrict__ bary, uint32_t n,
                                                        uint32_t num_tri, float4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 o0 = make_float4(0.f, 0.f, 0.f, 0.f), o1 = o0;
    const uint32_t p = prim[i];
    if (p < num_tri) {
        const float4 r1 = rays[2 * i + 1];
        const f3 rd = {r1.x, r1.y, r1.z};
        const float *t = tri + 9ull * p;
        const TriHit h{bary[3 * i], bary[3 * i + 1], bary[3 * i + 2], 0.f};
        const HitGeom g = hit_geom(h, load3(t), load3(t + 3), load3(t + 6));
        const f3 ff = dot3(g.n, neg3(rd)) < 0.0f ? neg3(g.n) : g.n;
        Pcg rng = pcg_seq(i);
        const float u0 = rng.uniform(), u1 = rng.uniform();
        // concentric disk (sampling.h:61-83), then z = sqrt(1 - x^2 - y^2) in FP64
        const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
        float dx = 0.f, dy = 0.f;
        if (!(ox == 0.0f && oy == 0.0f)) {
            const float pi = 3.14159274101257324f;
            const float pi_over_4 = pi / 4.0f;
            float rr, theta;
            if (fabsf(ox) > fabsf(oy)) {
                rr = ox;
                theta = pi_over_4 * (oy / ox);
            } else {
                rr = oy;
                theta = pi / 2.0f - pi_over_4 * (ox / oy);
            }
            // cos and sin in double, rounded once to float (the oracle does the same)
            dx = rr * float(cos(double(theta)));
            dy = rr * float(sin(double(theta)));
        }
        const float zz = float(1.0 - double(dx * dx) - double(dy * dy));
        const float z = sqrtf(zz > 0.0f ? zz : 0.0f);
        if (z != 0.0f) {
            f3 bx, by;
            basis(ff, &bx, &by);
            const f3 wi = add3(add3(mul3(dx, bx), mul3(dy, by)), mul3(z, ff));
            const f3 org = offset_origin(g, g.n, wi);
            o0 = make_float4(org.x, org.y, org.z, __builtin_inff());
            o1 = make_float4(wi.x, wi.y, wi.z, 0.f);
        }
    }
    out[2 * i] = o0;
    out[2 * i + 1] = o1;
}

}  // namespace

hipError_t trace_binary(const BvhTrace &B, const float *rays, uint32_t n, int any_hit, uint32_t *prim_out,
                        float *t_out, float *bary_out, uint32_t *err, unsigned long long *counts,
                        hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (n + kTraceBlock - 1) / kTraceBlock;
    const float4 *r = reinterpret_cast<const float4 *>(rays);
    if (counts)
        trace_binary_kernel<true><<<grid, kTraceBlock, 0, stream>>>(B.nodes, 2 * B.num_prims - 1, B.num_prims,
                                                                    B.tri_leaf, B.prim_order, r, n, any_hit,
                                                                    prim_out, t_out, bary_out, err, counts);
    else
        trace_binary_kernel<false><<<grid, kTraceBlock, 0, stream>>>(B.nodes, 2 * B.num_prims - 1, B.num_prims,
                                                                     B.tri_leaf, B.prim_order, r, n, any_hit,
                                                                     prim_out, t_out, bary_out, err, nullptr);
    return hipGetLastError();
}

hipError_t deindex_triangles(const float *verts, const uint32_t *idx, uint32_t nt, uint32_t nv, float *tri,
                             uint32_t *err, hipStream_t stream) {
    if (nt == 0) return hipSuccess;
    deindex_kernel<<<(nt + 255) / 256, 256, 0, stream>>>(verts, idx, nt, nv, tri, err);
    return hipGetLastError();
}

hipError_t ao_bounce(const float *tri, uint32_t num_tri, const float *rays, const uint32_t *prim, const float *bary,
                     uint32_t n, float *rays_out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    ao_bounce_kernel<<<(n + 255) / 256, 256, 0, stream>>>(tri, reinterpret_cast<const float4 *>(rays), prim, bary,
                                                          n, num_tri, reinterpret_cast<float4 *>(rays_out));
    return hipGetLastError();
}

}  // namespace pbrtm
