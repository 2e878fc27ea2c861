// C ABI of libpbrtm.so (include/pbrtm.h): contexts, aggregate handles, error
// reporting, host<->device staging and the deterministic workload generators.
#include "../../include/pbrtm.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "hlbvh_build.h"
#include "pathtrace.h"
#include "pbrtm_geometry.h"
#include "traverse.h"
#include "wide_bvh.h"

using namespace pbrtm;

// PBRTM_BATCH_AUTOTUNE state of one batch signature: the candidates tried and
// their measured times.
struct BatchTune {
    uint64_t key = 0;
    int tried = 0;
    uint32_t pool[6] = {}, quarters[6] = {};
    double ms[6] = {};
};

// Device buffers of the last aggregate freed on a context, handed to its next
// build (one set), so that rebuilding in a loop allocates nothing.
struct SpareBvh {
    DevBuf nodes, prim_order, tri_leaf, wnodes, worder;
};

// dst takes src's allocation (dst's own is released first).
static void take_buf(DevBuf &dst, DevBuf &src) {
    dst.release();
    dst.ptr = src.ptr;
    dst.bytes = src.bytes;
    src.ptr = nullptr;
    src.bytes = 0;
}

struct pbrtm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    BuildScratch scratch;
    DevBuf stage_in, stage_a, stage_b, stage_c, dev_err, spill, collapse, queue;
    pbrtm::pt::PathPool batch_pool;
    DevBuf batch_jobs, batch_films;
    hipEvent_t ev[8] = {};
    // a second stream (default priority) for build work that overlaps the wide
    // collapse: aux_ev[0] orders it after the binary build, aux_ev[1] the
    // collapse's end after it
    hipStream_t aux = nullptr;
    hipEvent_t aux_ev[2] = {};
    std::string error;
    std::vector<BatchTune> batch_tune;
    SpareBvh spare;
};

struct pbrtm_bvh {
    pbrtm_ctx *ctx = nullptr;  // must outlive the handle
    int device = 0;
    uint32_t n = 0;
    BvhDevice dev;
    WideBvh wide;
    DevBuf tri_owned, tri_leaf;
    const float *tri = nullptr;  // device, input order (owned or caller's)
    float bounds[6];
};

namespace pbrtm {

hipError_t DevBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return hipSuccess;
    release();
    size_t sz = std::max<size_t>(need, 256);
    hipError_t e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return e;
    }
    bytes = sz;
    return hipSuccess;
}

void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

std::string device_error_text(uint32_t bits) {
    std::string s;
    auto add = [&](uint32_t b, const char *t) {
        if (bits & b) {
            if (!s.empty()) s += "; ";
            s += t;
        }
    };
    add(kErrNonFinite, "non-finite vertex coordinate or vertex index out of range");
    add(kErrDegenerateCentroids, "treelet centroids coincide on the split axis (hlbvh.cu:670-674)");
    add(kErrSpin, "radix sort look-back did not converge");
    add(kErrStack, "traversal stack overflow");
    add(kErrSplit, "SAH split left one side empty");
    add(kErrCapacity, "build work list overflow");
    return s;
}

}  // namespace pbrtm

namespace {

int fail(pbrtm_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->error = msg;
    return code;
}

int hip_fail(pbrtm_ctx *ctx, hipError_t e, const char *where) {
    return fail(ctx, e == hipErrorOutOfMemory ? PBRTM_ERR_NOMEM : PBRTM_ERR_HIP,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define CK(ctx, expr, where)                           \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return hip_fail(ctx, _e, where); \
    } while (0)

// Host pointer -> device (staged in `buf`), or the device pointer as is.
hipError_t stage_in(pbrtm_ctx *ctx, DevBuf &buf, const void *p, size_t bytes, uint32_t flags, const void **out) {
    if (flags & PBRTM_DEVICE_PTRS) {
        *out = p;
        return hipSuccess;
    }
    hipError_t e = buf.reserve(bytes);
    if (e != hipSuccess) return e;
    if (bytes) e = hipMemcpyAsync(buf.ptr, p, bytes, hipMemcpyHostToDevice, ctx->stream);
    *out = buf.ptr;
    return e;
}

double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

// Scratch of pbrtm_bvh_intersect_noalloc: a file-scope device array, so that the
// entry point allocates nothing. Layout: a 256-byte header (word 0: device
// error bits, word 1: ray queue length; cleared by a kernel on every call),
// the ray queue (one word per ray), then the traversal's stack spill area.
// Each device holds its own copy for the life of the process; every call on a
// device shares it, so calls that overlap on one device race.
constexpr uint64_t kNoAllocMaxRays = 1ull << 22;
constexpr uint64_t kNoAllocHeader = 256;
constexpr uint64_t kNoAllocSpillBytes = 16ull << 20;
constexpr uint64_t kNoAllocBytes = kNoAllocHeader + kNoAllocMaxRays * 4 + kNoAllocSpillBytes;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

// The array's device address per device, taken once with that device current
// (by pbrtm_create, or by the first no-alloc call there).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<uint8_t *>(p);
    return hipSuccess;
}

// Device address of this device's scratch (the device must be current), or null.
uint8_t *noalloc_scratch_of(int device) {
    if (noalloc_scratch_init(device) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

__global__ void zero_words_kernel(uint32_t *w, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0u;
}

__global__ void fill_miss_kernel(uint32_t n, uint32_t *prim, float *t, float *bary) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    prim[i] = kNoHit;
    if (t) t[i] = 0.f;
    if (bary) bary[3 * size_t(i)] = bary[3 * size_t(i) + 1] = bary[3 * size_t(i) + 2] = 0.f;
}

uint32_t canon_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u == 0x80000000u ? 0u : u;
}

void put(std::vector<uint8_t> &out, uint32_t v) {
    uint8_t b[4];
    std::memcpy(b, &v, 4);
    out.insert(out.end(), b, b + 4);
}

}  // namespace

extern "C" {

int pbrtm_device_count(int *count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return PBRTM_OK;
}

int pbrtm_create(pbrtm_ctx **out, int device) {
    if (!out) return PBRTM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PBRTM_ERR_UNSUPPORTED;
    if (device < 0 || device >= count) return PBRTM_ERR_INVALID;
    auto *ctx = new (std::nothrow) pbrtm_ctx();
    if (!ctx) return PBRTM_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PBRTM_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return PBRTM_ERR_HIP;
    }
    for (auto &e : ctx->ev) hipEventCreate(&e);
    for (auto &e : ctx->aux_ev) hipEventCreateWithFlags(&e, hipEventDisableTiming);
    // the no-alloc entry point's scratch address on this device (checked again there)
    (void)noalloc_scratch_init(device);
    *out = ctx;
    return PBRTM_OK;
}

void pbrtm_destroy(pbrtm_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->scratch.release();
    for (DevBuf *b : {&ctx->spare.nodes, &ctx->spare.prim_order, &ctx->spare.tri_leaf, &ctx->spare.wnodes,
                      &ctx->spare.worder})
        b->release();
    ctx->stage_in.release();
    ctx->stage_a.release();
    ctx->stage_b.release();
    ctx->stage_c.release();
    ctx->dev_err.release();
    ctx->spill.release();
    ctx->collapse.release();
    ctx->queue.release();
    auto &p = ctx->batch_pool;
    for (DevBuf *b : {&p.path_id, &p.flags, &p.rng, &p.dim, &p.lam, &p.lam_pdf, &p.L, &p.beta, &p.nee_L, &p.scal,
                      &p.prev, &p.rays, &p.prim, &p.thit, &p.bary, &p.srays, &p.sprim, &p.rec, &p.rec_round,
                      &p.counters, &p.queue, &p.squeue, &p.spill, &p.job, &p.sround, &ctx->batch_jobs,
                      &ctx->batch_films, &p.rsort.keys[0], &p.rsort.keys[1], &p.rsort.vals[0], &p.rsort.vals[1],
                      &p.rsort.hist, &p.rsort.sort.status, &p.rsort.sort.offsets, &p.rsort.sort.counters})
        b->release();
    for (auto &e : ctx->ev) hipEventDestroy(e);
    for (auto &e : ctx->aux_ev) hipEventDestroy(e);
    hipStreamSynchronize(ctx->aux);
    hipStreamDestroy(ctx->aux);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *pbrtm_last_error(const pbrtm_ctx *ctx) { return ctx ? ctx->error.c_str() : "null context"; }

void *pbrtm_stream(pbrtm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int pbrtm_morton_sort(pbrtm_ctx *ctx, const float *tri, uint64_t n, uint32_t flags, uint32_t *codes_out,
                      uint32_t *order_out) {
    if (!ctx) return PBRTM_ERR_INVALID;
    if (n >= (1ull << 29)) return fail(ctx, PBRTM_ERR_INVALID, "at most 2^29 - 1 triangles");
    if (n == 0) return PBRTM_OK;
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const void *d_tri = nullptr;
    CK(ctx, stage_in(ctx, ctx->stage_in, tri, n * 36, flags, &d_tri), "upload triangles");
    uint32_t *K = nullptr, *V = nullptr;
    CK(ctx, morton_sort_device((const float *)d_tri, uint32_t(n), ctx->scratch, &K, &V, ctx->stream), "morton sort");
    const hipMemcpyKind kind = (flags & PBRTM_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    CK(ctx, hipMemcpyAsync(codes_out, K, n * 4, kind, ctx->stream), "copy codes");
    CK(ctx, hipMemcpyAsync(order_out, V, n * 4, kind, ctx->stream), "copy order");
    uint32_t err = 0;
    CK(ctx, hipMemcpyAsync(&err, ctx->scratch.counters.as<uint32_t>() + kCntErr, 4, hipMemcpyDeviceToHost,
                           ctx->stream), "copy error");
    CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (err) return fail(ctx, PBRTM_ERR_DEVICE, device_error_text(err));
    return PBRTM_OK;
}

int pbrtm_bvh_build(pbrtm_ctx *ctx, const float *verts, uint64_t nv, const uint32_t *idx, uint64_t nt,
                    uint32_t flags, pbrtm_bvh **out, pbrtm_build_stats *stats) {
    if (!ctx || !out) return PBRTM_ERR_INVALID;
    *out = nullptr;
    if (nt >= (1ull << 29)) return fail(ctx, PBRTM_ERR_INVALID, "at most 2^29 - 1 triangles");
    if (!idx && nv != 3 * nt) return fail(ctx, PBRTM_ERR_INVALID, "a soup needs nv == 3 * nt");
    if (idx && nv >= (1ull << 32)) return fail(ctx, PBRTM_ERR_INVALID, "too many vertices");
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const auto t0 = std::chrono::steady_clock::now();
    auto *bvh = new (std::nothrow) pbrtm_bvh();
    if (!bvh) return fail(ctx, PBRTM_ERR_NOMEM, "handle");
    bvh->ctx = ctx;
    bvh->device = ctx->device;
    bvh->n = uint32_t(nt);
    take_buf(bvh->dev.nodes, ctx->spare.nodes);  // kept if large enough, else reallocated
    take_buf(bvh->dev.prim_order, ctx->spare.prim_order);
    take_buf(bvh->tri_leaf, ctx->spare.tri_leaf);
    take_buf(bvh->wide.nodes, ctx->spare.wnodes);
    take_buf(bvh->wide.order, ctx->spare.worder);
    const float mx = 3.402823466e+38f;
    const float empty[6] = {mx, mx, mx, -mx, -mx, -mx};
    std::memcpy(bvh->bounds, empty, sizeof(empty));
    auto bail = [&](int code) {
        hipStreamSynchronize(ctx->aux);  // no leaf-order work left on buffers the context keeps
        pbrtm_bvh_free(bvh);
        return code;
    };
    if (nt > 0) {
        if (idx) {
            const void *dv = nullptr, *di = nullptr;
            hipError_t e = stage_in(ctx, ctx->stage_in, verts, nv * 12, flags, &dv);
            if (e == hipSuccess) e = stage_in(ctx, ctx->stage_a, idx, nt * 12, flags, &di);
            if (e == hipSuccess) e = bvh->tri_owned.reserve(nt * 36);
            if (e == hipSuccess) e = ctx->dev_err.reserve(4);
            if (e == hipSuccess) e = hipMemsetAsync(ctx->dev_err.ptr, 0, 4, ctx->stream);
            if (e == hipSuccess)
                e = deindex_triangles((const float *)dv, (const uint32_t *)di, uint32_t(nt), uint32_t(nv),
                                      bvh->tri_owned.as<float>(), ctx->dev_err.as<uint32_t>(), ctx->stream);
            uint32_t derr = 0;
            if (e == hipSuccess)
                e = hipMemcpyAsync(&derr, ctx->dev_err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return bail(hip_fail(ctx, e, "mesh upload"));
            if (derr) return bail(fail(ctx, PBRTM_ERR_INVALID, "vertex index out of range"));
            bvh->tri = bvh->tri_owned.as<float>();
        } else if (flags & PBRTM_DEVICE_PTRS) {
            bvh->tri = verts;
        } else {
            hipError_t e = bvh->tri_owned.reserve(nt * 36);
            if (e == hipSuccess) e = hipMemcpyAsync(bvh->tri_owned.ptr, verts, nt * 36, hipMemcpyHostToDevice, ctx->stream);
            if (e != hipSuccess) return bail(hip_fail(ctx, e, "triangle upload"));
            bvh->tri = bvh->tri_owned.as<float>();
        }
        BuildPhaseTimes times;
        for (int k = 0; k < 6; ++k) times.ev[k] = ctx->ev[k];
        uint32_t derr = 0;
        hipError_t e = build_hlbvh_device(bvh->tri, uint32_t(nt), ctx->scratch, bvh->dev, &derr, ctx->stream, &times);
        if (e != hipSuccess) return bail(hip_fail(ctx, e, "build"));
        if (derr) return bail(fail(ctx, PBRTM_ERR_DEVICE, device_error_text(derr)));
        e = bvh->tri_leaf.reserve(nt * 48);
        // the leaf order streams on the second stream while the collapse's first
        // levels (a few nodes each, bound by their dependent fetches) run here
        if (e == hipSuccess) e = hipEventRecord(ctx->aux_ev[0], ctx->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->aux, ctx->aux_ev[0], 0);
        if (e == hipSuccess) e = leaf_order(ctx->scratch, bvh->dev, bvh->tri_leaf.as<float>(), ctx->aux);
        if (e == hipSuccess) e = hipEventRecord(ctx->aux_ev[1], ctx->aux);
        if (e == hipSuccess) e = ctx->dev_err.reserve(64);
        if (e == hipSuccess) e = hipMemsetAsync(ctx->dev_err.ptr, 0, 4, ctx->stream);
        if (e == hipSuccess)
            e = collapse_wide(bvh->dev.nodes.as<BNode>(), uint32_t(nt), bvh->wide, ctx->collapse,
                              ctx->dev_err.as<uint32_t>(), ctx->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->aux_ev[1], 0);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev[6], ctx->stream);
        uint32_t cerr = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&cerr, ctx->dev_err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
        BNode root;
        if (e == hipSuccess) e = hipMemcpyAsync(&root, bvh->dev.nodes.ptr, sizeof(BNode), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return bail(hip_fail(ctx, e, "layout"));
        if (cerr) return bail(fail(ctx, PBRTM_ERR_DEVICE, device_error_text(cerr)));
        for (int k = 0; k < 3; ++k) {
            bvh->bounds[k] = root.lo[k];
            bvh->bounds[3 + k] = root.hi[k];
        }
        if (stats) {
            stats->ms_sort = event_ms(ctx->ev[0], ctx->ev[1]);
            stats->ms_treelets = event_ms(ctx->ev[1], ctx->ev[2]);
            stats->ms_top = event_ms(ctx->ev[2], ctx->ev[3]);
            stats->ms_bottom = event_ms(ctx->ev[3], ctx->ev[4]);
            stats->ms_layout = event_ms(ctx->ev[5], ctx->ev[6]);
        }
    }
    if (stats) {
        if (nt == 0) {
            stats->ms_sort = stats->ms_treelets = stats->ms_top = stats->ms_bottom = stats->ms_layout = 0.0;
        }
        stats->num_prims = nt;
        stats->num_treelets = bvh->dev.num_treelets;
        stats->top_levels = bvh->dev.top_levels;
        stats->bottom_levels = bvh->dev.bottom_levels;
        stats->small_top = bvh->dev.small_top;
        stats->small_bottom = bvh->dev.small_bottom;
        stats->wide_nodes = bvh->wide.num_nodes;
        stats->wide_depth = bvh->wide.max_depth;
        stats->ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    *out = bvh;
    return PBRTM_OK;
}

void pbrtm_bvh_free(pbrtm_bvh *bvh) {
    if (!bvh) return;
    hipSetDevice(bvh->device);  // hipFree waits for work that uses the memory
    if (bvh->ctx) {  // the context's next build reuses these (its work is ordered after this one's on the stream)
        SpareBvh &s = bvh->ctx->spare;
        take_buf(s.nodes, bvh->dev.nodes);
        take_buf(s.prim_order, bvh->dev.prim_order);
        take_buf(s.tri_leaf, bvh->tri_leaf);
        take_buf(s.wnodes, bvh->wide.nodes);
        take_buf(s.worder, bvh->wide.order);
    }
    bvh->dev.nodes.release();
    bvh->dev.prim_order.release();
    bvh->tri_owned.release();
    bvh->tri_leaf.release();
    bvh->wide.nodes.release();
    bvh->wide.order.release();
    delete bvh;
}

int pbrtm_bvh_bounds(const pbrtm_bvh *bvh, float out6[6]) {
    if (!bvh || !out6) return PBRTM_ERR_INVALID;
    std::memcpy(out6, bvh->bounds, sizeof(bvh->bounds));
    return PBRTM_OK;
}

int pbrtm_bvh_export_canonical(const pbrtm_bvh *bvh, uint8_t *buf, uint64_t *len) {
    if (!bvh || !len) return PBRTM_ERR_INVALID;
    pbrtm_ctx *ctx = bvh->ctx;
    std::vector<uint8_t> out;
    const uint32_t n = bvh->n;
    if (n > 0) {
        CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
        std::vector<BNode> nodes(size_t(2) * n - 1);
        std::vector<uint32_t> order(n);
        CK(ctx, hipMemcpyAsync(nodes.data(), bvh->dev.nodes.ptr, nodes.size() * sizeof(BNode),
                               hipMemcpyDeviceToHost, ctx->stream), "copy nodes");
        CK(ctx, hipMemcpyAsync(order.data(), bvh->dev.prim_order.ptr, size_t(n) * 4, hipMemcpyDeviceToHost,
                               ctx->stream), "copy order");
        CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
        struct Frame {
            uint32_t id, prims;
        };
        std::vector<Frame> stack{{0u, n}};
        std::vector<uint32_t> ids;
        while (!stack.empty()) {
            Frame f = stack.back();
            stack.pop_back();
            if (f.id >= nodes.size()) return fail(ctx, PBRTM_ERR_DEVICE, "node id out of range");
            const BNode &nd = nodes[f.id];
            const bool leaf = (nd.a & kLeafFlag) != 0;
            put(out, leaf ? 1u : ((nd.b >> 30) << 8));
            for (int k = 0; k < 3; ++k) put(out, canon_bits(nd.lo[k]));
            for (int k = 0; k < 3; ++k) put(out, canon_bits(nd.hi[k]));
            put(out, f.prims);
            if (leaf) {
                const uint32_t first = nd.a & ~kLeafFlag;
                if (nd.b != f.prims || uint64_t(first) + nd.b > n)
                    return fail(ctx, PBRTM_ERR_DEVICE, "leaf range inconsistent");
                ids.assign(order.begin() + first, order.begin() + first + nd.b);
                std::sort(ids.begin(), ids.end());
                for (uint32_t v : ids) put(out, v);
                continue;
            }
            const uint32_t right = nd.b & 0x3FFFFFFFu;
            const uint32_t lp = (right - f.id) / 2;
            if (nd.a != f.id + 1 || lp == 0 || lp >= f.prims)
                return fail(ctx, PBRTM_ERR_DEVICE, "interior node inconsistent");
            stack.push_back({right, f.prims - lp});
            stack.push_back({nd.a, lp});
        }
    }
    if (!buf) {
        *len = out.size();
        return PBRTM_OK;
    }
    if (*len < out.size()) return fail(ctx, PBRTM_ERR_INVALID, "buffer too small");
    std::memcpy(buf, out.data(), out.size());
    *len = out.size();
    return PBRTM_OK;
}

int pbrtm_bvh_intersect(pbrtm_bvh *bvh, const float *rays, uint64_t n, int any_hit, uint32_t flags,
                        uint32_t *prim_out, float *t_out, float *bary_out, uint64_t *stats_out) {
    if (!bvh || !prim_out) return PBRTM_ERR_INVALID;
    pbrtm_ctx *ctx = bvh->ctx;
    if (n >= (1ull << 32)) return fail(ctx, PBRTM_ERR_INVALID, "too many rays");
    if (n == 0) return PBRTM_OK;
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const bool dev = flags & PBRTM_DEVICE_PTRS;
    const void *d_rays = nullptr;
    CK(ctx, stage_in(ctx, ctx->stage_in, rays, n * 32, flags, &d_rays), "upload rays");
    uint32_t *d_prim = prim_out;
    float *d_t = t_out, *d_b = bary_out;
    if (!dev) {
        CK(ctx, ctx->stage_a.reserve(n * 4), "alloc");
        CK(ctx, ctx->stage_b.reserve(n * 4), "alloc");
        CK(ctx, ctx->stage_c.reserve(n * 12), "alloc");
        d_prim = ctx->stage_a.as<uint32_t>();
        d_t = t_out ? ctx->stage_b.as<float>() : nullptr;
        d_b = bary_out ? ctx->stage_c.as<float>() : nullptr;
    }
    CK(ctx, ctx->dev_err.reserve(64), "alloc");
    CK(ctx, hipMemsetAsync(ctx->dev_err.ptr, 0, 64, ctx->stream), "memset");
    unsigned long long *d_counts = stats_out ? (unsigned long long *)(ctx->dev_err.as<uint8_t>() + 16) : nullptr;
    const bool binary = (flags & PBRTM_TRACE_BINARY) != 0;
    if (bvh->n == 0) {
        CK(ctx, hipMemsetAsync(d_prim, 0xFF, n * 4, ctx->stream), "memset");
        if (d_t) CK(ctx, hipMemsetAsync(d_t, 0, n * 4, ctx->stream), "memset");
        if (d_b) CK(ctx, hipMemsetAsync(d_b, 0, n * 12, ctx->stream), "memset");
    } else {
        if (binary) {
            BvhTrace B{bvh->dev.nodes.as<BNode>(), bvh->tri_leaf.as<float>(), bvh->dev.prim_order.as<uint32_t>(),
                       bvh->n};
            CK(ctx, trace_binary(B, (const float *)d_rays, uint32_t(n), any_hit, d_prim, d_t, d_b,
                                 ctx->dev_err.as<uint32_t>(), d_counts, ctx->stream), "trace");
        } else {
            WideTrace W{};
            W.nodes = bvh->wide.nodes.as<WNode>();
            W.order = bvh->wide.order.as<uint32_t>();
            W.tri_leaf = bvh->tri_leaf.as<float>();
            W.prim_order = bvh->dev.prim_order.as<uint32_t>();
            W.num_nodes = bvh->wide.num_nodes;
            W.num_prims = bvh->n;
            W.max_depth = bvh->wide.max_depth;
            W.spill_per_ray = W.max_depth > kLdsStack ? W.max_depth - kLdsStack + 1 : 0;
            if (W.spill_per_ray) {
                CK(ctx, ctx->spill.reserve(size_t(n) * W.spill_per_ray * sizeof(uint2)), "alloc stack spill");
                W.spill = ctx->spill.as<uint2>();
            }
            CK(ctx, ctx->queue.reserve(size_t(n) * 4), "alloc ray queue");
            CK(ctx, trace_wide(W, (const float *)d_rays, uint32_t(n), any_hit, d_prim, d_t, d_b,
                               ctx->dev_err.as<uint32_t>(), d_counts, ctx->queue.as<uint32_t>(),
                               ctx->dev_err.as<uint32_t>() + 1, ctx->stream), "trace");
        }
    }
    if (!dev) {
        CK(ctx, hipMemcpyAsync(prim_out, d_prim, n * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
        if (t_out) CK(ctx, hipMemcpyAsync(t_out, d_t, n * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
        if (bary_out) CK(ctx, hipMemcpyAsync(bary_out, d_b, n * 12, hipMemcpyDeviceToHost, ctx->stream), "download");
    }
    uint32_t derr = 0;
    CK(ctx, hipMemcpyAsync(&derr, ctx->dev_err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    if (stats_out)
        CK(ctx, hipMemcpyAsync(stats_out, d_counts, 24, hipMemcpyDeviceToHost, ctx->stream), "download");
    CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (derr) return fail(ctx, PBRTM_ERR_DEVICE, device_error_text(derr));
    return PBRTM_OK;
}

int pbrtm_bvh_intersect_noalloc(pbrtm_bvh *bvh, const float *rays, uint64_t n, int any_hit, uint32_t *prim_out,
                                float *t_out, float *bary_out) {
    if (!bvh || !rays || !prim_out) return PBRTM_ERR_INVALID;
    pbrtm_ctx *ctx = bvh->ctx;
    if (n > kNoAllocMaxRays)
        return fail(ctx, PBRTM_ERR_INVALID,
                    "pbrtm_bvh_intersect_noalloc: at most " + std::to_string(kNoAllocMaxRays) + " rays per call");
    if (n == 0) return PBRTM_OK;
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    uint8_t *base = noalloc_scratch_of(ctx->device);
    if (!base) return fail(ctx, PBRTM_ERR_HIP, "pbrtm_bvh_intersect_noalloc: scratch symbol address unavailable");
    uint32_t *words = reinterpret_cast<uint32_t *>(base);  // [0] device error bits, [1] queue length
    uint32_t *queue = reinterpret_cast<uint32_t *>(base + kNoAllocHeader);
    uint2 *spill = reinterpret_cast<uint2 *>(base + kNoAllocHeader + kNoAllocMaxRays * sizeof(uint32_t));
    zero_words_kernel<<<1, 64, 0, ctx->stream>>>(words, kNoAllocHeader / sizeof(uint32_t));
    CK(ctx, hipGetLastError(), "clear scratch");
    if (bvh->n == 0) {
        fill_miss_kernel<<<uint32_t((n + 255) / 256), 256, 0, ctx->stream>>>(uint32_t(n), prim_out, t_out, bary_out);
        CK(ctx, hipGetLastError(), "fill misses");
    } else {
        WideTrace W{};
        W.nodes = bvh->wide.nodes.as<WNode>();
        W.order = bvh->wide.order.as<uint32_t>();
        W.tri_leaf = bvh->tri_leaf.as<float>();
        W.prim_order = bvh->dev.prim_order.as<uint32_t>();
        W.num_nodes = bvh->wide.num_nodes;
        W.num_prims = bvh->n;
        W.max_depth = bvh->wide.max_depth;
        W.spill_per_ray = W.max_depth > kLdsStack ? W.max_depth - kLdsStack + 1 : 0;
        if (uint64_t(W.spill_per_ray) * n * sizeof(uint2) > kNoAllocSpillBytes)
            return fail(ctx, PBRTM_ERR_INVALID,
                        "pbrtm_bvh_intersect_noalloc: this tree (" + std::to_string(W.max_depth) +
                            " wide levels) needs a stack spill area larger than the " +
                            std::to_string(kNoAllocSpillBytes) + "-byte scratch for " + std::to_string(n) + " rays");
        W.spill = W.spill_per_ray ? spill : nullptr;
        CK(ctx, trace_wide(W, rays, uint32_t(n), any_hit, prim_out, t_out, bary_out, words, nullptr, queue,
                           words + 1, ctx->stream), "trace");
    }
    uint32_t derr = 0;
    CK(ctx, hipMemcpyAsync(&derr, words, 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (derr) return fail(ctx, PBRTM_ERR_DEVICE, device_error_text(derr));
    return PBRTM_OK;
}

int pbrtm_ao_bounce(pbrtm_bvh *bvh, const float *rays, const uint32_t *prim, const float *bary, uint64_t n,
                    uint32_t flags, float *rays_out) {
    if (!bvh || !rays || !prim || !bary || !rays_out) return PBRTM_ERR_INVALID;
    pbrtm_ctx *ctx = bvh->ctx;
    if (n >= (1ull << 32)) return fail(ctx, PBRTM_ERR_INVALID, "too many rays");
    if (n == 0) return PBRTM_OK;
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    for (uint64_t i = 0; !(flags & PBRTM_DEVICE_PTRS) && i < n; ++i)
        if (prim[i] != PBRTM_NO_HIT && prim[i] >= bvh->n) return fail(ctx, PBRTM_ERR_INVALID, "prim id out of range");
    const void *d_rays = nullptr, *d_prim = nullptr, *d_bary = nullptr;
    CK(ctx, stage_in(ctx, ctx->stage_in, rays, n * 32, flags, &d_rays), "upload");
    CK(ctx, stage_in(ctx, ctx->stage_a, prim, n * 4, flags, &d_prim), "upload");
    CK(ctx, stage_in(ctx, ctx->stage_b, bary, n * 12, flags, &d_bary), "upload");
    float *d_out = rays_out;
    if (!(flags & PBRTM_DEVICE_PTRS)) {
        CK(ctx, ctx->stage_c.reserve(n * 32), "alloc");
        d_out = ctx->stage_c.as<float>();
    }
    CK(ctx, ao_bounce(bvh->tri, bvh->n, (const float *)d_rays, (const uint32_t *)d_prim, (const float *)d_bary,
                      uint32_t(n), d_out, ctx->stream), "bounce");
    if (!(flags & PBRTM_DEVICE_PTRS))
        CK(ctx, hipMemcpyAsync(rays_out, d_out, n * 32, hipMemcpyDeviceToHost, ctx->stream), "download");
    CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return PBRTM_OK;
}

int pbrtm_gen_soup(uint64_t seed, uint64_t n, float mu, float sigma, int threads, float *tri_out) {
    if (!tri_out && n) return PBRTM_ERR_INVALID;
    const Pcg base = pcg_seq(seed);
    auto work = [&](uint64_t b, uint64_t e) {
        const double two_pi = 6.283185307179586;
        for (uint64_t i = b; i < e; ++i) {
            // advance the stream to draw 16*i (Brown's log-time jump)
            uint64_t cur_mult = 0x5851f42d4c957f2dULL, cur_plus = base.inc, acc_mult = 1u, acc_plus = 0u;
            for (uint64_t delta = 16 * i; delta > 0; delta >>= 1) {
                if (delta & 1) {
                    acc_mult *= cur_mult;
                    acc_plus = acc_plus * cur_mult + cur_plus;
                }
                cur_plus = (cur_mult + 1) * cur_plus;
                cur_mult *= cur_mult;
            }
            Pcg r{acc_mult * base.state + acc_plus, base.inc};
            const double cx = r.uniform(), cy = r.uniform(), cz = r.uniform();
            float *t = tri_out + 9 * i;
            for (int v = 0; v < 3; ++v) {
                const double ua = r.uniform(), ub = r.uniform(), uc = r.uniform(), ud = r.uniform();
                const double g = std::sqrt(-2.0 * std::log(1.0 - ua)) * std::cos(two_pi * ub);
                const double len = std::exp(double(mu) + double(sigma) * g);
                const double z = 1.0 - 2.0 * uc;
                const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
                const double phi = two_pi * ud;
                t[3 * v] = float(cx + len * rr * std::cos(phi));
                t[3 * v + 1] = float(cy + len * rr * std::sin(phi));
                t[3 * v + 2] = float(cz + len * z);
            }
        }
    };
    threads = std::max(1, std::min(threads, 64));
    if (threads == 1 || n < 65536) {
        work(0, n);
        return PBRTM_OK;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + threads - 1) / threads;
    for (int k = 0; k < threads; ++k) {
        const uint64_t b = std::min(n, k * chunk), e = std::min(n, b + chunk);
        pool.emplace_back(work, b, e);
    }
    for (auto &t : pool) t.join();
    return PBRTM_OK;
}

int pbrtm_gen_camera_rays(uint32_t width, uint32_t height, const float eye[3], const float look[3],
                          const float up[3], float fov_deg, float *rays_out) {
    if (!eye || !look || !up || !rays_out || width == 0 || height == 0) return PBRTM_ERR_INVALID;
    auto norm = [](double v[3]) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= l; v[1] /= l; v[2] /= l;
    };
    auto crs = [](const double a[3], const double b[3], double o[3]) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double f[3] = {double(look[0]) - eye[0], double(look[1]) - eye[1], double(look[2]) - eye[2]};
    norm(f);
    double u0[3] = {up[0], up[1], up[2]}, r[3], u[3];
    crs(u0, f, r);
    norm(r);
    crs(f, r, u);
    const double th = std::tan(double(fov_deg) * 3.141592653589793 / 360.0);
    const double aspect = double(width) / double(height);
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            const double sx = (2.0 * (x + 0.5) / width - 1.0) * th * aspect;
            const double sy = (1.0 - 2.0 * (y + 0.5) / height) * th;
            double d[3] = {f[0] + sx * r[0] + sy * u[0], f[1] + sx * r[1] + sy * u[1], f[2] + sx * r[2] + sy * u[2]};
            norm(d);
            float *o = rays_out + 8 * (uint64_t(y) * width + x);
            o[0] = eye[0]; o[1] = eye[1]; o[2] = eye[2];
            o[3] = INFINITY;
            o[4] = float(d[0]); o[5] = float(d[1]); o[6] = float(d[2]);
            o[7] = 0.f;
        }
    }
    return PBRTM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ path tracer

struct pbrtm_scene {
    pbrtm_ctx *ctx = nullptr;  // must outlive the handle
    int device = 0;
    pbrtm_bvh *bvh = nullptr;
    pbrtm::pt::SceneHost host;
    pbrtm::DevBuf tri, tri_info, mats, lights, pmf, cdf, spec_lam, spec_val, filt, shapes, prim_shape;
    pbrtm::pt::SceneDev dev{};
    pbrtm::pt::PathPool pool;
    pbrtm::DevBuf film_rgb, film_aux;
    pbrtm::DevBuf g_lfin, g_pfin, g_target, g_adj, g_coef, g_acc, g_slot;  // gradient replay
};

extern "C" {

void pbrtm_scene_free(pbrtm_scene *sc) {
    if (!sc) return;
    hipSetDevice(sc->device);
    if (sc->bvh) pbrtm_bvh_free(sc->bvh);
    for (pbrtm::DevBuf *b : {&sc->tri, &sc->tri_info, &sc->mats, &sc->lights, &sc->pmf, &sc->cdf, &sc->spec_lam,
                             &sc->spec_val, &sc->filt, &sc->shapes, &sc->prim_shape, &sc->film_rgb, &sc->film_aux,
                             &sc->g_lfin, &sc->g_pfin,
                             &sc->g_target, &sc->g_adj, &sc->g_coef, &sc->g_acc, &sc->g_slot})
        b->release();
    auto &p = sc->pool;
    for (pbrtm::DevBuf *b : {&p.path_id, &p.flags, &p.rng, &p.dim, &p.lam, &p.lam_pdf, &p.L, &p.beta, &p.nee_L,
                             &p.scal, &p.prev, &p.rays, &p.prim, &p.thit, &p.bary, &p.srays, &p.sprim, &p.rec,
                             &p.rec_round, &p.counters, &p.queue, &p.squeue, &p.spill, &p.job, &p.sround,
                             &p.rsort.keys[0], &p.rsort.keys[1], &p.rsort.vals[0], &p.rsort.vals[1], &p.rsort.hist,
                             &p.rsort.sort.status, &p.rsort.sort.offsets, &p.rsort.sort.counters})
        b->release();
    delete sc;
}

int pbrtm_scene_create(pbrtm_ctx *ctx, const pbrtm_scene_desc *desc, pbrtm_scene **out) {
    if (!ctx || !desc || !out) return PBRTM_ERR_INVALID;
    *out = nullptr;
    if (desc->num_tris == 0) return fail(ctx, PBRTM_ERR_INVALID, "scene has no triangles");
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    auto *sc = new (std::nothrow) pbrtm_scene();
    if (!sc) return fail(ctx, PBRTM_ERR_NOMEM, "handle");
    sc->ctx = ctx;
    sc->device = ctx->device;
    auto bail = [&](int code) {
        pbrtm_scene_free(sc);
        return code;
    };
    const std::string msg = pbrtm::pt::build_scene_host(desc, sc->host);
    if (!msg.empty()) return bail(fail(ctx, PBRTM_ERR_INVALID, msg));
    const auto &H = sc->host;
    auto up = [&](pbrtm::DevBuf &b, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = b.reserve(std::max<size_t>(bytes, 4));
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream);
        return e;
    };
    hipError_t e = up(sc->tri, H.tri.data(), H.tri.size() * 4);
    if (e == hipSuccess) e = up(sc->tri_info, H.tri_info.data(), H.tri_info.size() * sizeof(uint2));
    if (e == hipSuccess) e = up(sc->mats, H.mats.data(), H.mats.size() * sizeof(pbrtm::pt::DMat));
    if (e == hipSuccess) e = up(sc->lights, H.lights.data(), H.lights.size() * sizeof(pbrtm::pt::DLight));
    if (e == hipSuccess) e = up(sc->pmf, H.light_pmf.data(), H.light_pmf.size() * 4);
    if (e == hipSuccess) e = up(sc->cdf, H.light_cdf.data(), H.light_cdf.size() * 4);
    if (e == hipSuccess) e = up(sc->spec_lam, H.spec_lam.data(), H.spec_lam.size() * 4);
    if (e == hipSuccess) e = up(sc->spec_val, H.spec_val.data(), H.spec_val.size() * 4);
    const size_t fn = H.filt_n;
    std::vector<float> filt;
    for (const auto *v : {&H.filt_f, &H.filt_cond_func, &H.filt_cond_cdf, &H.filt_cond_int, &H.filt_marg_func,
                          &H.filt_marg_cdf})
        filt.insert(filt.end(), v->begin(), v->end());
    if (e == hipSuccess) e = up(sc->filt, filt.data(), filt.size() * 4);
    if (e == hipSuccess && !H.shapes.empty()) e = up(sc->shapes, H.shapes.data(), H.shapes.size() * sizeof(pbrtm::pt::DShape));
    if (e == hipSuccess && !H.prim_shape.empty()) e = up(sc->prim_shape, H.prim_shape.data(), H.prim_shape.size() * 4);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return bail(hip_fail(ctx, e, "scene upload"));
    pbrtm::pt::SceneDev &D = sc->dev;
    D.tri = sc->tri.as<float>();
    D.tri_info = sc->tri_info.as<uint2>();
    D.mats = sc->mats.as<pbrtm::pt::DMat>();
    D.lights = sc->lights.as<pbrtm::pt::DLight>();
    D.light_pmf = sc->pmf.as<float>();
    D.light_cdf = sc->cdf.as<float>();
    D.num_tris = uint32_t(desc->num_tris);
    D.num_lights = uint32_t(H.lights.size());
    D.spec_lam = sc->spec_lam.as<float>();
    D.spec_val = sc->spec_val.as<float>();
    for (int k = 0; k < 3; ++k) D.cie_off[k] = H.cie_off[k];
    D.cie_n = H.cie_n;
    D.d65_off = H.d65_off;
    D.d65_n = H.d65_n;
    D.cie_uni = H.cie_uni;
    D.d65_uni = H.d65_uni;
    D.cie_shared = H.cie_shared;
    D.cie_inv_step = H.cie_inv_step;
    D.d65_inv_step = H.d65_inv_step;
    std::memcpy(D.cam_from_raster, H.cam_from_raster, sizeof(D.cam_from_raster));
    std::memcpy(D.render_from_camera, H.render_from_camera, sizeof(D.render_from_camera));
    std::memcpy(D.cam_origin, H.cam_origin, sizeof(D.cam_origin));
    const float *fb = sc->filt.as<float>();
    D.filt_f = fb;
    D.filt_cond_func = fb + fn * fn;
    D.filt_cond_cdf = D.filt_cond_func + fn * fn;
    D.filt_cond_int = D.filt_cond_cdf + fn * (fn + 1);
    D.filt_marg_func = D.filt_cond_int + fn;
    D.filt_marg_cdf = D.filt_marg_func + fn;
    D.filt_marg_int = H.filt_marg_int;
    D.filt_radius = H.filt_radius;
    D.filt_n = H.filt_n;
    std::memcpy(D.out_from_sensor, H.out_from_sensor, sizeof(D.out_from_sensor));
    D.width = H.width;
    D.height = H.height;
    D.ao_scale = H.ao_scale;
    std::memcpy(D.sn_c, H.sn_c, sizeof(D.sn_c));
    D.num_inf = H.num_inf;
    std::memcpy(D.inf_idx, H.inf_idx, sizeof(D.inf_idx));
    D.inf_dist = H.inf_dist;
    D.shapes = H.shapes.empty() ? nullptr : sc->shapes.as<pbrtm::pt::DShape>();
    D.prim_shape = H.shapes.empty() ? nullptr : sc->prim_shape.as<uint32_t>();
    // The aggregate over the render-space primitives (SceneBuilder::preprocess,
    // scene_builder.cu:829-838): triangles, and analytic primitives through their boxes.
    const int rc = pbrtm_bvh_build(ctx, sc->tri.as<float>(), 3 * desc->num_tris, nullptr, desc->num_tris,
                                   PBRTM_DEVICE_PTRS, &sc->bvh, nullptr);
    if (rc != PBRTM_OK) return bail(rc);
    if (D.shapes) {
        e = pbrtm::pt::tag_leaf_shapes(sc->bvh->tri_leaf.as<float>(), sc->bvh->dev.prim_order.as<uint32_t>(),
                                       D.prim_shape, uint32_t(desc->num_tris), ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return bail(hip_fail(ctx, e, "shape tags"));
    }
    *out = sc;
    return PBRTM_OK;
}

}  // extern "C"

namespace {

WideTrace scene_trace(const pbrtm_scene *sc) {
    const pbrtm_bvh *bvh = sc->bvh;
    WideTrace W{};
    W.nodes = bvh->wide.nodes.as<WNode>();
    W.order = bvh->wide.order.as<uint32_t>();
    W.tri_leaf = bvh->tri_leaf.as<float>();
    W.prim_order = bvh->dev.prim_order.as<uint32_t>();
    W.num_nodes = bvh->wide.num_nodes;
    W.num_prims = bvh->n;
    W.max_depth = bvh->wide.max_depth;
    W.shapes = sc->dev.shapes;
    W.spill_per_ray = W.max_depth > kLdsStack ? W.max_depth - kLdsStack + 1 : 0;
    for (int a = 0; a < 3; ++a) {
        W.lo[a] = bvh->bounds[a];
        W.hi[a] = bvh->bounds[3 + a];
    }
    return W;
}

// Shared checks of pbrtm_render and pbrtm_render_batch; an empty string when valid.
uint32_t effective_pool(const pbrtm_render_params *params, uint64_t wh) {
    if (params->pool_size) return params->pool_size;
    const uint32_t integ = params->flags & PBRTM_INTEGRATOR_MASK;
    if (integ == PBRTM_INTEGRATOR_PATH) return 2u << 20;
    const uint64_t want = integ == PBRTM_INTEGRATOR_MEGAKERNEL_PATH ? (4ull << 20) : (8ull << 20);
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(want, wh * params->spp)));
}

const char *check_render_params(const pbrtm_scene *sc, const pbrtm_render_params *params, int *code) {
    const uint32_t spp = params->spp;
    const uint32_t spd = uint32_t(std::sqrt(double(spp)));
    *code = PBRTM_ERR_INVALID;
    if (spp == 0 || spd * spd != spp) return "spp must be a positive square (stratified)";
    *code = PBRTM_ERR_UNSUPPORTED;
    if (spp > 1024) return "at most 1024 samples per pixel";
    const uint64_t wh = uint64_t(sc->host.width) * sc->host.height;
    if (wh * spp + effective_pool(params, wh) >= 0xFFFFFFFFull)
        return "at most 2^32 - 1 - pool_size paths per render";
    *code = PBRTM_ERR_INVALID;
    if (params->max_depth > 0xFFFE) return "max_depth too large";
    if ((params->flags & PBRTM_INTEGRATOR_MASK) > PBRTM_INTEGRATOR_SURFACE_NORMAL) return "unknown integrator";
    return "";
}

}  // namespace

extern "C" {

int pbrtm_integrator_from_name(const char *name, uint32_t *flags_out) {
    if (!name || !flags_out) return PBRTM_ERR_INVALID;
    static const struct {
        const char *name;
        uint32_t flags;
    } kNames[] = {{"path", PBRTM_INTEGRATOR_PATH},
                  {"volpath", PBRTM_INTEGRATOR_PATH},
                  {"megakernelpath", PBRTM_INTEGRATOR_MEGAKERNEL_PATH},
                  {"ambientocclusion", PBRTM_INTEGRATOR_AMBIENT_OCCLUSION},
                  {"surfacenormal", PBRTM_INTEGRATOR_SURFACE_NORMAL}};
    for (const auto &k : kNames)
        if (std::strcmp(name, k.name) == 0) {
            *flags_out = k.flags;
            return PBRTM_OK;
        }
    return PBRTM_ERR_UNSUPPORTED;
}

int pbrtm_render(pbrtm_ctx *ctx, pbrtm_scene *sc, const pbrtm_render_params *params, float *film_rgb,
                 float *film_aux, pbrtm_render_stats *stats) {
    if (!ctx || !sc || !params || !film_rgb) return PBRTM_ERR_INVALID;
    if (sc->ctx != ctx) return fail(ctx, PBRTM_ERR_INVALID, "scene belongs to another context");
    int code = 0;
    const char *why = check_render_params(sc, params, &code);
    if (*why) return fail(ctx, code, why);
    const uint32_t spp = params->spp;
    const uint64_t wh = uint64_t(sc->host.width) * sc->host.height;
    // Default pool: the reference's PATH_POOL_SIZE for the wavefront integrator (its
    // films depend on it); the megakernel integrators' films do not (samples in
    // sample order), and their measured optimum on a 1080p 16 spp render is 2^22
    // slots for megakernelpath and 2^23 for the single-vertex ones, capped at the
    // render's path count.
    const uint32_t pool = effective_pool(params, wh);
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const bool dev = params->flags & PBRTM_DEVICE_PTRS;
    float *d_rgb = film_rgb, *d_aux = film_aux;
    if (!dev) {
        CK(ctx, sc->film_rgb.reserve(wh * 12), "alloc film");
        d_rgb = sc->film_rgb.as<float>();
        if (film_aux) {
            CK(ctx, sc->film_aux.reserve(wh * 16), "alloc film");
            d_aux = sc->film_aux.as<float>();
        }
    }
    const WideTrace W = scene_trace(sc);
    const bool count = (params->flags & PBRTM_RENDER_COUNT) != 0;
    CK(ctx, ctx->dev_err.reserve(128), "alloc");
    CK(ctx, hipMemsetAsync(ctx->dev_err.ptr, 0, count ? 128 : 4, ctx->stream), "memset");
    unsigned long long *d_counts = count ? reinterpret_cast<unsigned long long *>(ctx->dev_err.as<uint8_t>() + 64)
                                         : nullptr;
    pbrtm::pt::RenderStats rs;
    CK(ctx, pbrtm::pt::render(sc->dev, W, spp, params->max_depth, pool, sc->pool, d_rgb, d_aux,
                              ctx->dev_err.as<uint32_t>(), &rs, ctx->stream, nullptr, nullptr, nullptr,
                              (params->flags & PBRTM_INTEGRATOR_MASK) >> 4, d_counts),
       "render");
    uint32_t derr = 0;
    unsigned long long hcounts[6] = {0, 0, 0, 0, 0, 0};
    CK(ctx, hipMemcpyAsync(&derr, ctx->dev_err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    if (count) CK(ctx, hipMemcpyAsync(hcounts, d_counts, sizeof(hcounts), hipMemcpyDeviceToHost, ctx->stream), "download");
    if (!dev) {
        CK(ctx, hipMemcpyAsync(film_rgb, d_rgb, wh * 12, hipMemcpyDeviceToHost, ctx->stream), "download");
        if (film_aux) CK(ctx, hipMemcpyAsync(film_aux, d_aux, wh * 16, hipMemcpyDeviceToHost, ctx->stream), "download");
    }
    CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (derr) return fail(ctx, PBRTM_ERR_DEVICE, device_error_text(derr));
    if (stats) {
        stats->ms_total = rs.ms_total;
        stats->ms_trace = rs.ms_trace;
        stats->ms_shadow = rs.ms_shadow;
        stats->ms_shade = rs.ms_shade;
        stats->ms_film = rs.ms_film;
        stats->rounds = rs.rounds;
        stats->rays = rs.rays;
        stats->shadow_rays = rs.shadow_rays;
        stats->paths = rs.paths;
        stats->ray_order = rs.sorted;
        stats->grid_quarters = rs.grid_quarters;
        for (int k = 0; k < 3; ++k) {
            stats->closest_counts[k] = hcounts[k];
            stats->shadow_counts[k] = hcounts[3 + k];
        }
    }
    return PBRTM_OK;
}

int pbrtm_render_grad(pbrtm_ctx *ctx, pbrtm_scene *sc, const pbrtm_render_params *params, const float *target_rgb,
                      float *film_rgb, double *loss, float *d_albedo, float *d_scale, pbrtm_grad_stats *stats) {
    if (!ctx || !sc || !params || !target_rgb || !loss) return PBRTM_ERR_INVALID;
    if (sc->ctx != ctx) return fail(ctx, PBRTM_ERR_INVALID, "scene belongs to another context");
    int code = 0;
    const char *why = check_render_params(sc, params, &code);
    if (*why) return fail(ctx, code, why);
    if (params->flags & PBRTM_INTEGRATOR_MASK)
        return fail(ctx, PBRTM_ERR_UNSUPPORTED, "gradients render with the wavefront path integrator only");
    if (sc->host.num_inf)
        return fail(ctx, PBRTM_ERR_UNSUPPORTED, "gradients cover area lights only (the scene has an infinite light)");
    const pbrtm::pt::SceneHost &H = sc->host;
    const uint32_t ng = uint32_t(H.group_scale.size());
    if (ng > 255) return fail(ctx, PBRTM_ERR_UNSUPPORTED, "gradients support at most 255 emitters");
    const uint32_t nm = uint32_t(H.mats.size());
    const uint32_t spp = params->spp;
    const uint64_t wh = uint64_t(H.width) * H.height, total = wh * spp;
    const uint32_t pool = params->pool_size ? params->pool_size : (2u << 20);
    const auto t0 = std::chrono::steady_clock::now();
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    CK(ctx, sc->g_lfin.reserve(total * 16), "alloc");
    CK(ctx, sc->g_pfin.reserve(total * 16), "alloc");
    CK(ctx, sc->film_rgb.reserve(wh * 12), "alloc");
    CK(ctx, sc->film_aux.reserve(wh * 16), "alloc");
    CK(ctx, sc->g_target.reserve(wh * 12), "alloc");
    CK(ctx, sc->g_adj.reserve(wh * 12), "alloc");
    const size_t coef_bytes = size_t(std::max(nm, 1u)) * 12 * pbrtm::pt::kGradShards;
    CK(ctx, sc->g_coef.reserve(coef_bytes), "alloc");
    CK(ctx, sc->g_acc.reserve(size_t(ng + 2) * 8), "alloc");
    CK(ctx, sc->g_slot.reserve(size_t(pool) * 48), "alloc");
    CK(ctx, ctx->dev_err.reserve(64), "alloc");
    hipStream_t st = ctx->stream;
    CK(ctx, hipMemsetAsync(ctx->dev_err.ptr, 0, 4, st), "memset");
    CK(ctx, hipMemcpyAsync(sc->g_target.ptr, target_rgb, wh * 12, hipMemcpyHostToDevice, st), "upload");
    CK(ctx, hipMemsetAsync(sc->g_coef.ptr, 0, coef_bytes, st), "memset");
    CK(ctx, hipMemsetAsync(sc->g_acc.ptr, 0, size_t(ng + 2) * 8, st), "memset");
    const WideTrace W = scene_trace(sc);
    float *film = sc->film_rgb.as<float>(), *aux = sc->film_aux.as<float>();
    float4 *lfin = sc->g_lfin.as<float4>(), *pfin = sc->g_pfin.as<float4>();
    uint32_t *err = ctx->dev_err.as<uint32_t>();
    pbrtm::pt::RenderStats fwd, rep;
    CK(ctx, pbrtm::pt::render(sc->dev, W, spp, params->max_depth, pool, sc->pool, film, aux, err, &fwd, st, lfin, pfin),
       "render");
    const auto t1 = std::chrono::steady_clock::now();
    double *acc = sc->g_acc.as<double>();
    CK(ctx, pbrtm::pt::film_adjoint(sc->dev, film, aux, sc->g_target.as<float>(), sc->g_adj.as<float>(), acc + ng + 1,
                                    st), "adjoint");
    CK(ctx, hipStreamSynchronize(st), "sync");
    const auto t2 = std::chrono::steady_clock::now();
    pbrtm::pt::GradDev gd{};
    gd.lfin = lfin;
    gd.pfin = pfin;
    gd.adj = sc->g_adj.as<float>();
    gd.d_coef = sc->g_coef.as<float>();
    gd.ncoef = 3 * nm;
    gd.acc = acc;
    gd.g = sc->g_slot.as<float4>();
    gd.ltot = gd.g + pool;
    gd.nee = gd.g + 2 * size_t(pool);
    gd.ngroups = ng;
    CK(ctx, pbrtm::pt::render(sc->dev, W, spp, params->max_depth, pool, sc->pool, nullptr, nullptr, err, &rep, st,
                              nullptr, nullptr, &gd), "replay");
    const auto t3 = std::chrono::steady_clock::now();
    std::vector<float> coef(size_t(nm) * 3);
    std::vector<double> hacc(ng + 2);
    uint32_t derr = 0;
    CK(ctx, hipMemcpyAsync(&derr, err, 4, hipMemcpyDeviceToHost, st), "download");
    if (nm) CK(ctx, hipMemcpyAsync(coef.data(), gd.d_coef, coef.size() * 4, hipMemcpyDeviceToHost, st), "download");
    CK(ctx, hipMemcpyAsync(hacc.data(), acc, hacc.size() * 8, hipMemcpyDeviceToHost, st), "download");
    if (film_rgb) CK(ctx, hipMemcpyAsync(film_rgb, film, wh * 12, hipMemcpyDeviceToHost, st), "download");
    CK(ctx, hipStreamSynchronize(st), "sync");
    if (derr) return fail(ctx, PBRTM_ERR_DEVICE, device_error_text(derr));
    *loss = hacc[ng + 1];
    if (d_albedo) {  // J^T through the RGB-to-spectrum table (non-diffuse materials accumulate nothing)
        if (nm) pbrtm::pt::rgb_to_sigmoid_jacobian_transpose(H.mat_rgb.data(), coef.data(), nm, d_albedo);
        for (uint32_t m = 0; m < nm; ++m)
            if (H.mats[m].type != pbrtm::pt::kDiffuse) d_albedo[3 * size_t(m)] = d_albedo[3 * size_t(m) + 1] =
                d_albedo[3 * size_t(m) + 2] = 0.f;
    }
    if (d_scale) {  // dLoss/ds_m = (a_m - k P_m) / s_m, P_m = the emitter's total pmf
        std::vector<double> P(ng, 0.0);
        for (size_t t = 0; t < H.lights.size(); ++t) P[H.lights[t].group] += double(H.light_pmf[t]);
        for (uint32_t m = 0; m < ng; ++m)
            d_scale[m] = H.group_scale[m] != 0.f ? float((hacc[m] - hacc[ng] * P[m]) / double(H.group_scale[m])) : 0.f;
    }
    if (stats) {
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        stats->ms_forward = ms(t0, t1);
        stats->ms_adjoint = ms(t1, t2);
        stats->ms_replay = ms(t2, t3);
        stats->ms_total = ms(t0, std::chrono::steady_clock::now());
        stats->rounds = fwd.rounds;
        stats->rays = fwd.rays;
        stats->shadow_rays = fwd.shadow_rays;
        stats->paths = fwd.paths;
    }
    return PBRTM_OK;
}

int pbrtm_render_batch(pbrtm_ctx *ctx, pbrtm_job *jobs, uint32_t njobs, const pbrtm_batch_params *params,
                       pbrtm_batch_stats *stats) {
    if (!ctx || !jobs || !params || njobs == 0) return PBRTM_ERR_INVALID;
    const bool dev = params->flags & PBRTM_DEVICE_PTRS;
    uint64_t film_floats = 0, paths = 0, shared_paths = 0;
    const uint64_t solo =
        (params->solo_paths == 0 || params->solo_paths == 0xFFFFFFFFu) ? ~uint64_t(0) : params->solo_paths;
    for (uint32_t j = 0; j < njobs; ++j) {
        const pbrtm_job &jb = jobs[j];
        if (!jb.scene || !jb.film_rgb) return fail(ctx, PBRTM_ERR_INVALID, "job without scene or film");
        if (jb.scene->ctx != ctx) return fail(ctx, PBRTM_ERR_INVALID, "scene belongs to another context");
        int code = 0;
        const char *why = check_render_params(jb.scene, &jb.params, &code);
        if (*why) return fail(ctx, code, why);
        if (jb.params.flags & PBRTM_INTEGRATOR_MASK)
            return fail(ctx, PBRTM_ERR_UNSUPPORTED, "batched jobs render with the wavefront path integrator only");
        if (jb.params.max_depth > 4000) return fail(ctx, PBRTM_ERR_UNSUPPORTED, "batched jobs allow max_depth <= 4000");
        const uint64_t wh = uint64_t(jb.scene->host.width) * jb.scene->host.height;
        film_floats += wh * (jb.film_aux ? 7 : 3);
        paths += wh * jb.params.spp;
        if (wh * jb.params.spp < solo) shared_paths += wh * jb.params.spp;  // the jobs of the shared pool
    }
    // PBRTM_BATCH_AUTOTUNE: this batch's tuning record (FNV-1a over its jobs)
    BatchTune *tune = nullptr;
    uint32_t quarters = 0, tune_state = 0, tune_pool = 0;
    if (params->flags & PBRTM_BATCH_AUTOTUNE) {
        uint64_t key = 1469598103934665603ull;
        auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
        mix(njobs);
        mix(solo);
        for (uint32_t j = 0; j < njobs; ++j) {
            mix(reinterpret_cast<uintptr_t>(jobs[j].scene));
            mix(jobs[j].params.spp);
            mix(jobs[j].params.max_depth);
            mix(jobs[j].params.pool_size);
        }
        for (auto &t : ctx->batch_tune)
            if (t.key == key) tune = &t;
        if (!tune) {
            ctx->batch_tune.emplace_back();
            tune = &ctx->batch_tune.back();
            tune->key = key;
        }
        const uint32_t pools[3] = {8u << 20, 16u << 20, 32u << 20};
        auto best_of = [&](int upto) {
            int b = 0;
            for (int k = 1; k < upto; ++k)
                if (tune->ms[k] < tune->ms[b]) b = k;
            return b;
        };
        if (tune->tried < 3) {  // pool sizes at the full grid
            tune_pool = pools[tune->tried];
            quarters = 4;
            tune_state = 1;
        } else if (tune->tried < 5) {  // the best pool at 3 and 2 quarters of the grid
            tune_pool = tune->pool[best_of(3)];
            quarters = tune->tried == 3 ? 3u : 2u;
            tune_state = 1;
        } else {
            const int b = best_of(5);
            tune_pool = tune->pool[b];
            quarters = tune->quarters[b];
            tune_state = 2;
        }
    }
    // the shared pool's slots: pool_size, or 2^24, capped at its jobs' path count (as render_batch does)
    const uint32_t pool = tune ? tune_pool : (params->pool_size ? params->pool_size : (16u << 20));
    const uint64_t slots = std::min<uint64_t>(pool, shared_paths);
    if (shared_paths + slots >= 0xFFFFFFFFull || paths >= 0xFFFFFFFFull)
        return fail(ctx, PBRTM_ERR_UNSUPPORTED, "at most 2^32 - 1 - pool_size paths per batch");
    CK(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (!dev) CK(ctx, ctx->batch_films.reserve(film_floats * 4), "alloc films");
    std::vector<pbrtm::pt::BatchJob> bj(njobs);
    float *f = dev ? nullptr : ctx->batch_films.as<float>();
    for (uint32_t j = 0; j < njobs; ++j) {
        const pbrtm_job &jb = jobs[j];
        const uint64_t wh = uint64_t(jb.scene->host.width) * jb.scene->host.height;
        bj[j].S = jb.scene->dev;
        bj[j].W = scene_trace(jb.scene);
        bj[j].spp = jb.params.spp;
        bj[j].max_depth = jb.params.max_depth;
        bj[j].vpool = jb.params.pool_size ? jb.params.pool_size : (2u << 20);
        if (dev) {
            bj[j].film_rgb = jb.film_rgb;
            bj[j].film_aux = jb.film_aux;
        } else {
            bj[j].film_rgb = f;
            f += wh * 3;
            bj[j].film_aux = jb.film_aux ? f : nullptr;
            if (jb.film_aux) f += wh * 4;
        }
    }
    const uint32_t rpl = params->rounds_per_launch ? params->rounds_per_launch : 8u;
    CK(ctx, ctx->dev_err.reserve(64), "alloc");
    CK(ctx, hipMemsetAsync(ctx->dev_err.ptr, 0, 4, ctx->stream), "memset");
    pbrtm::pt::BatchStats bs;
    CK(ctx, pbrtm::pt::render_batch(bj, pool, rpl, !(params->flags & PBRTM_BATCH_NO_GRAPH), solo, ctx->batch_pool,
                                    ctx->batch_jobs, ctx->dev_err.as<uint32_t>(), &bs, ctx->stream, quarters),
       "render batch");
    uint32_t derr = 0;
    CK(ctx, hipMemcpyAsync(&derr, ctx->dev_err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    if (!dev) {
        for (uint32_t j = 0; j < njobs; ++j) {
            const uint64_t wh = uint64_t(jobs[j].scene->host.width) * jobs[j].scene->host.height;
            CK(ctx, hipMemcpyAsync(jobs[j].film_rgb, bj[j].film_rgb, wh * 12, hipMemcpyDeviceToHost, ctx->stream),
               "download");
            if (jobs[j].film_aux)
                CK(ctx, hipMemcpyAsync(jobs[j].film_aux, bj[j].film_aux, wh * 16, hipMemcpyDeviceToHost, ctx->stream),
                   "download");
        }
    }
    CK(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (derr) return fail(ctx, PBRTM_ERR_DEVICE, device_error_text(derr));
    for (uint32_t j = 0; j < njobs; ++j) jobs[j].rounds = bs.job_rounds[j];
    if (tune && tune_state == 1) {
        const int k = tune->tried++;
        tune->pool[k] = pool;
        tune->quarters[k] = quarters;
        tune->ms[k] = bs.ms_total;
    }
    if (stats) {
        stats->ms_total = bs.ms_total;
        stats->ms_rounds = bs.ms_rounds;
        stats->ms_schedule = bs.ms_schedule;
        stats->ms_film = bs.ms_film;
        stats->rounds = bs.rounds;
        stats->rays = bs.rays;
        stats->shadow_rays = bs.shadow_rays;
        stats->paths = bs.paths;
        stats->pool_size = uint32_t(slots);  // 0 when every job rendered alone
        stats->launches = bs.launches;
        stats->solo_jobs = bs.solo;
        stats->grid_quarters = quarters ? quarters : 4u;
        stats->tune_state = tune_state;
        stats->reserved = 0;
    }
    return PBRTM_OK;
}

int pbrtm_rgb_to_sigmoid(const float rgb[3], float out3[3]) {
    if (!rgb || !out3) return PBRTM_ERR_INVALID;
    try {
        pbrtm::pt::rgb_to_sigmoid(rgb, out3);
    } catch (const std::exception &) {
        return PBRTM_ERR_INVALID;
    }
    return PBRTM_OK;
}

int pbrtm_srgb_matrices(float xyz_from_rgb9[9], float rgb_from_xyz9[9], float white_xy2[2]) {
    if (!xyz_from_rgb9 || !rgb_from_xyz9 || !white_xy2) return PBRTM_ERR_INVALID;
    pbrtm::pt::srgb_matrices(xyz_from_rgb9, rgb_from_xyz9, white_xy2);
    return PBRTM_OK;
}

}  // extern "C"
