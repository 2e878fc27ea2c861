// TEST INFRASTRUCTURE ONLY — CPU oracle for the HLBVH build and traversal.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker or the timed CPU baseline.
//
// A host-only restatement of accelerator/hlbvh.cu (reference), FP32 without
// contraction. Each function cites the reference lines it follows.
//
// The reference numbers nodes and orders primitives inside a treelet through
// atomics (hlbvh.cu:65, :167, :796), so only its tree of primitive SETS is
// deterministic. This oracle builds that set-tree and exports it in a canonical
// preorder form; the GPU build must reproduce the same bytes.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"
#include "oracle_math.h"
#include "oracle_rng.h"
#include "oracle_tri.h"

using namespace orc;

namespace {

constexpr uint32_t kTreeletShift = 9;   // hlbvh.cu:7-9: 30 - 21 low bits dropped
constexpr uint32_t kNumBuckets = 24;    // hlbvh.cu:34
constexpr int kStackCap = 128;          // hlbvh.cu:193 Stack<uint, 128>

inline V3 tri_point(const float *tri, uint64_t i, int v) {
    const float *p = tri + 9 * i + 3 * v;
    return {p[0], p[1], p[2]};
}

struct Prim {          // hlbvh.h:14-19 MortonPrimitive
    uint32_t idx;
    uint32_t code;
    B3 bounds;
    V3 centroid;
};

struct Node {          // hlbvh.h:33-74, with explicit children
    B3 bounds;
    int axis;          // -1 = leaf
    int left, right;
    uint32_t first, count;  // leaf range in prim order; subtree prim count
};

// Node storage: ids are handed out by an atomic counter, as the reference's
// node_count.fetch_add(2) (hlbvh.cu:796), so concurrent subtree builds can
// allocate; a build over n primitives uses at most 2n - 1 nodes.
struct NodeArena {
    std::unique_ptr<Node[]> p;
    uint64_t cap = 0;
    std::atomic<uint64_t> used{0};
    void reset(uint64_t c) {
        p.reset(new Node[c]);  // trivially constructible: the pages are touched when used
        cap = c;
        used = 0;
    }
    Node &operator[](uint64_t i) { return p[i]; }
    const Node &operator[](uint64_t i) const { return p[i]; }
    uint64_t size() const { return used.load(); }
    bool empty() const { return size() == 0; }
};

struct Treelet {       // hlbvh.h:21-25
    uint32_t first, n;
    B3 bounds;
};

void parallel_for(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)> &fn);

}  // namespace

struct orc_bvh {
    const float *tri = nullptr;  // caller-owned, must outlive the handle
    uint64_t n = 0;
    orc_prim_test custom_test = nullptr;  // primitives with custom[i] != 0 are tested by this
    void *custom_user = nullptr;
    const uint8_t *custom = nullptr;
    std::vector<Prim> prims;     // Morton-sorted, then partitioned in place
    NodeArena nodes;
    int top_nodes = 0;
    std::string error;
};

namespace {

void parallel_for(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)> &fn) {
    if (threads <= 1 || n < 4096) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<uint64_t> next{0};
    const uint64_t chunk = 1024;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                uint64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    }
    for (auto &th : pool) th.join();
}

// hlbvh.cu:101-141 and the host centroid-bounds loop :428-446. With threads > 1
// the centroid bounds are reduced per chunk; min and max do not depend on the
// order (up to the sign of a zero, which reaches no Morton code).
void compute_morton(const float *tri, uint64_t n, std::vector<Prim> &prims, int threads = 1) {
    prims.resize(n);
    B3 cb = b3_empty();
    std::mutex mu;
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        B3 part = b3_empty();
        for (uint64_t i = b; i < e; ++i) {
            Prim &p = prims[i];
            p.idx = uint32_t(i);
            p.bounds = b3_of_tri(tri_point(tri, i, 0), tri_point(tri, i, 1), tri_point(tri, i, 2));
            p.centroid = b3_centroid(p.bounds);
            part = b3_add_point(part, p.centroid);
        }
        std::lock_guard<std::mutex> g(mu);
        cb = b3_add_point(b3_add_point(cb, part.lo), part.hi);
    });
    int max_dim = b3_max_dimension(cb);
    float radius = (cb.hi[max_dim] - cb.lo[max_dim]) / 2;
    V3 lo = cb.lo, hi = cb.hi;
    for (int d = 0; d < 3; ++d) {
        if (d == max_dim) continue;
        float center = b3_centroid(cb)[d];
        lo[d] = center - radius;
        hi[d] = center + radius;
    }
    B3 adj{vmin(lo, hi), vmax(lo, hi)};
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            V3 o = b3_offset(adj, prims[i].centroid);
            V3 s = scale(float(1 << 10), o);
            prims[i].code = encode_morton3(f2u_sat(s.x), f2u_sat(s.y), f2u_sat(s.z));
        }
    });
}

// Canonical order: full 30-bit code, ties by input index. It refines the
// reference's treelet-only counting sort (hlbvh.cu:54-80, :457-502). Keys are
// (code << 32 | index), distributed over 1024 buckets by the top 10 code bits,
// each bucket sorted on its own; prims[i].idx == i on entry.
void canonical_sort(std::vector<Prim> &prims, int threads = 1) {
    const uint64_t n = prims.size();
    constexpr int kBits = 10, kBuckets = 1 << kBits;
    const int parts = std::max(1, threads);
    std::vector<uint64_t> hist(uint64_t(parts) * kBuckets, 0);
    auto range = [&](int t) { return std::make_pair(n * t / parts, n * (t + 1) / parts); };
    auto bucket = [](uint32_t code) { return code >> (30 - kBits); };
    auto run = [&](const std::function<void(int)> &fn) {
        std::vector<std::thread> pool;
        for (int t = 1; t < parts; ++t) pool.emplace_back(fn, t);
        fn(0);
        for (auto &th : pool) th.join();
    };
    run([&](int t) {
        auto [b, e] = range(t);
        for (uint64_t i = b; i < e; ++i) ++hist[uint64_t(t) * kBuckets + bucket(prims[i].code)];
    });
    std::vector<uint64_t> start(kBuckets + 1, 0);
    uint64_t acc = 0;
    for (int k = 0; k < kBuckets; ++k) {  // bucket-major, part-minor offsets
        start[k] = acc;
        for (int t = 0; t < parts; ++t) {
            const uint64_t c = hist[uint64_t(t) * kBuckets + k];
            hist[uint64_t(t) * kBuckets + k] = acc;
            acc += c;
        }
    }
    start[kBuckets] = acc;
    std::unique_ptr<uint64_t[]> keys(new uint64_t[n]);
    run([&](int t) {
        auto [b, e] = range(t);
        uint64_t *off = hist.data() + uint64_t(t) * kBuckets;
        for (uint64_t i = b; i < e; ++i)
            keys[off[bucket(prims[i].code)]++] = (uint64_t(prims[i].code) << 32) | prims[i].idx;
    });
    std::atomic<int> next{0};
    run([&](int) {
        for (int k; (k = next.fetch_add(1)) < kBuckets;) std::sort(keys.get() + start[k], keys.get() + start[k + 1]);
    });
    std::vector<Prim> sorted(n);
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) sorted[i] = prims[uint32_t(keys[i])];
    });
    prims.swap(sorted);
}

struct Builder {
    orc_bvh *bvh;
    std::vector<Treelet> treelets;
    std::atomic<bool> failed{false};
    std::string why;
    std::mutex why_mu;

    void fail(const char *msg) {
        std::lock_guard<std::mutex> g(why_mu);
        if (!failed.exchange(true)) why = msg;
    }

    // k consecutive ids (the two children of a split), fields initialised as a leaf
    int new_nodes(int k) {
        const uint64_t id = bvh->nodes.used.fetch_add(uint64_t(k));
        if (id + k > bvh->nodes.cap) {  // cannot happen: a binary tree over n primitives has < 2n nodes
            fail("node arena exhausted");
            return 0;
        }
        for (int j = 0; j < k; ++j) bvh->nodes[id + j] = Node{b3_empty(), -1, -1, -1, 0u, 0u};
        return int(id);
    }

    // hlbvh.cu:649-826 — SAH over treelets, 24 buckets. spawn > 0 lets the left
    // subtree go to a thread of its own, as the reference hands subtrees of 64
    // or more treelets to its thread pool (hlbvh.cu:802-824).
    void build_upper(int node, const std::vector<uint32_t> &ids, int spawn = 0) {
        if (failed) return;
        if (ids.size() == 1) {
            const Treelet &t = treelets[ids[0]];
            Node &nd = bvh->nodes[node];
            nd.bounds = t.bounds;
            nd.axis = -1;
            nd.first = t.first;
            nd.count = t.n;
            return;
        }
        B3 full = b3_empty(), cb = b3_empty();
        uint32_t total = 0;
        for (uint32_t id : ids) {
            cb = b3_add_point(cb, b3_centroid(treelets[id].bounds));
            full = b3_union(full, treelets[id].bounds);
            total += treelets[id].n;
        }
        const int axis = b3_max_dimension(cb);
        if (cb.lo[axis] == cb.hi[axis]) {  // hlbvh.cu:670-674 (fatal in the reference)
            fail("degenerate treelet centroids");
            return;
        }
        struct Bucket {
            uint32_t count = 0;
            B3 bounds = b3_empty();
            std::vector<uint32_t> ids;
        } buckets[kNumBuckets];
        const float base = cb.lo[axis];
        const float span = cb.hi[axis] - cb.lo[axis];
        for (uint32_t id : ids) {
            float c = b3_centroid(treelets[id].bounds)[axis];
            uint32_t b = f2u_sat(float(kNumBuckets) * ((c - base) / span));
            if (b > kNumBuckets) {  // hlbvh.cu:696-698
                fail("bucket index out of range");
                return;
            }
            if (b == kNumBuckets) b = kNumBuckets - 1;
            buckets[b].count += treelets[id].n;
            buckets[b].bounds = b3_union(buckets[b].bounds, treelets[id].bounds);
            buckets[b].ids.push_back(id);
        }
        const float total_sa = b3_surface_area(full);
        float cost[kNumBuckets - 1];
        for (uint32_t s = 0; s + 1 < kNumBuckets; ++s) {  // hlbvh.cu:711-731
            B3 bl = b3_empty(), br = b3_empty();
            uint32_t cl = 0, cr = 0;
            for (uint32_t k = 0; k <= s; ++k) {
                bl = b3_union(bl, buckets[k].bounds);
                cl += buckets[k].count;
            }
            for (uint32_t k = s + 1; k < kNumBuckets; ++k) {
                br = b3_union(br, buckets[k].bounds);
                cr += buckets[k].count;
            }
            cost[s] = 0.125 + (float(cl) * b3_surface_area(bl) + float(cr) * b3_surface_area(br)) /
                                  total_sa;
        }
        float best = cost[0];
        uint32_t split = 0;
        for (uint32_t s = 1; s + 1 < kNumBuckets; ++s) {  // hlbvh.cu:734-741: first strict min
            if (cost[s] < best) {
                best = cost[s];
                split = s;
            }
        }
        std::vector<uint32_t> left, right;
        for (uint32_t k = 0; k <= split; ++k)
            left.insert(left.end(), buckets[k].ids.begin(), buckets[k].ids.end());
        for (uint32_t k = split + 1; k < kNumBuckets; ++k)
            right.insert(right.end(), buckets[k].ids.begin(), buckets[k].ids.end());
        if (left.empty() || right.empty()) {  // hlbvh.cu:758-777 midpoint fallback
            left.clear();
            right.clear();
            float mid = b3_centroid(cb)[axis];
            for (uint32_t id : ids) {
                if (b3_centroid(treelets[id].bounds)[axis] < mid) left.push_back(id);
                else right.push_back(id);
            }
            if (left.empty() || right.empty()) {
                fail("treelet split failed");
                return;
            }
        }
        const int l = new_nodes(2);
        const int r = l + 1;
        if (failed) return;
        Node &nd = bvh->nodes[node];
        nd.bounds = full;
        nd.axis = axis;
        nd.left = l;
        nd.right = r;
        nd.count = total;
        if (spawn > 0 && left.size() >= 64 && right.size() >= 64) {
            std::thread th([&]() { build_upper(l, left, spawn - 1); });
            build_upper(r, right, spawn - 1);
            th.join();
        } else {
            build_upper(l, left, spawn - 1);
            build_upper(r, right, spawn - 1);
        }
    }

    // hlbvh.cu:828-857 — Hoare partition on centroid[dim] < split.
    uint32_t partition(uint32_t start, uint32_t end, int dim, float split) {
        auto &P = bvh->prims;
        uint32_t left = start, right = end - 1;
        for (;;) {
            while (P[right].centroid[dim] >= split && right > start) right--;
            while (P[left].centroid[dim] < split && left < end - 1) left++;
            if (left < right) {
                std::swap(P[left], P[right]);
                continue;
            }
            if (left == start && right == start) return start;
            return right + 1;
        }
    }

    // hlbvh.cu:291-382 — midpoint split of a multi-primitive leaf, recursively
    // (the reference does one tree level per launch; the set-tree is the same).
    void build_bottom(int node) {
        Node nd = bvh->nodes[node];
        if (nd.axis >= 0 || nd.count <= 1) return;
        auto &P = bvh->prims;
        B3 cb = b3_empty();
        for (uint32_t i = nd.first; i < nd.first + nd.count; ++i) cb = b3_add_point(cb, P[i].centroid);
        int dim = b3_max_dimension(cb);
        float split = b3_centroid(cb)[dim];
        uint32_t mid = partition(nd.first, nd.first + nd.count, dim, split);
        if (mid == nd.first || mid == nd.first + nd.count) return;  // stays a leaf
        B3 lb = b3_empty(), rb = b3_empty();
        for (uint32_t i = nd.first; i < mid; ++i) lb = b3_union(lb, P[i].bounds);
        for (uint32_t i = mid; i < nd.first + nd.count; ++i) rb = b3_union(rb, P[i].bounds);
        const int l = new_nodes(2);
        const int r = l + 1;
        if (failed) return;
        Node &L = bvh->nodes[l];
        L.bounds = lb; L.axis = -1; L.first = nd.first; L.count = mid - nd.first;
        Node &R = bvh->nodes[r];
        R.bounds = rb; R.axis = -1; R.first = mid; R.count = nd.first + nd.count - mid;
        Node &me = bvh->nodes[node];
        me.axis = dim;
        me.left = l;
        me.right = r;
        me.bounds = b3_union(lb, rb);
        build_bottom(l);
        build_bottom(r);
    }
};

// -0.0 and +0.0 compare equal and the reference's min/max keeps whichever came
// first, so bound bit patterns are exported with -0.0 folded to +0.0.
uint32_t canon_bits(float f) {
    uint32_t u = f2u(f);
    return u == 0x80000000u ? 0u : u;
}

void put_u32(std::vector<uint8_t> &out, uint32_t v) {
    uint8_t b[4];
    std::memcpy(b, &v, 4);
    out.insert(out.end(), b, b + 4);
}

void export_node(const orc_bvh *bvh, int id, std::vector<uint8_t> &out, std::vector<uint32_t> &tmp) {
    const Node &nd = bvh->nodes[id];
    const bool leaf = nd.axis < 0;
    put_u32(out, leaf ? 1u : (0u | (uint32_t(nd.axis) << 8)));
    for (int d = 0; d < 3; ++d) put_u32(out, canon_bits(nd.bounds.lo[d]));
    for (int d = 0; d < 3; ++d) put_u32(out, canon_bits(nd.bounds.hi[d]));
    put_u32(out, nd.count);
    if (leaf) {
        tmp.clear();
        for (uint32_t i = nd.first; i < nd.first + nd.count; ++i) tmp.push_back(bvh->prims[i].idx);
        std::sort(tmp.begin(), tmp.end());
        for (uint32_t v : tmp) put_u32(out, v);
        return;
    }
    export_node(bvh, nd.left, out, tmp);
    export_node(bvh, nd.right, out, tmp);
}

struct TraceStats {
    uint64_t nodes = 0, tris = 0;
};

// hlbvh.cu:231-289 (closest) and :179-229 (any hit), over the oracle's tree.
// Returns -1 on stack overflow (fatal in the reference, stack.h:22-26).
int trace(const orc_bvh *bvh, V3 o, V3 d, float t_max, bool any_hit, uint32_t *prim, TriHit *hit,
          TraceStats *st) {
    *prim = 0xFFFFFFFFu;
    if (bvh->nodes.empty()) return 0;
    V3 inv = {float(1.0 / d.x), float(1.0 / d.y), float(1.0 / d.z)};
    int neg[3] = {int(inv.x < 0.0), int(inv.y < 0.0), int(inv.z < 0.0)};
    float best_t = t_max;
    int stack[kStackCap];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Node &nd = bvh->nodes[stack[--sp]];
        st->nodes++;
        if (!b3_fast_intersect(nd.bounds, o, best_t, inv, neg)) continue;
        if (nd.axis < 0) {
            for (uint32_t i = nd.first; i < nd.first + nd.count; ++i) {
                uint32_t pid = bvh->prims[i].idx;
                TriHit h;
                st->tris++;
                if (bvh->custom && bvh->custom[pid]) {
                    float t, b3[3];
                    if (!bvh->custom_test(bvh->custom_user, pid, &o.x, &d.x, best_t, &t, b3)) continue;
                    h = TriHit{b3[0], b3[1], b3[2], t};
                } else if (!intersect_triangle(o, d, best_t, tri_point(bvh->tri, pid, 0), tri_point(bvh->tri, pid, 1),
                                               tri_point(bvh->tri, pid, 2), &h)) {
                    continue;
                }
                *prim = pid;
                *hit = h;
                if (any_hit) return 1;
                best_t = h.t;
            }
            continue;
        }
        if (sp + 2 > kStackCap) return -1;
        if (neg[nd.axis]) {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        } else {
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
    return *prim != 0xFFFFFFFFu ? 1 : 0;
}

}  // namespace

extern "C" {

int orc_morton_codes(const float *tri, uint64_t n, uint32_t *codes_by_input) {
    std::vector<Prim> prims;
    compute_morton(tri, n, prims);
    for (uint64_t i = 0; i < n; ++i) codes_by_input[i] = prims[i].code;
    return 0;
}

int orc_morton_sort(const float *tri, uint64_t n, uint32_t *codes_out, uint32_t *order_out) {
    std::vector<Prim> prims;
    compute_morton(tri, n, prims);
    canonical_sort(prims);
    for (uint64_t i = 0; i < n; ++i) {
        codes_out[i] = prims[i].code;
        order_out[i] = prims[i].idx;
    }
    return 0;
}

// hlbvh.cu:384-624 — the whole build. tri is [n][3][3] floats, kept by pointer.
// threads > 1 splits the Morton pass, the sort, the top-level SAH (subtrees on
// threads of their own, as the reference's ThreadPool, hlbvh.cu:559-641) and the
// per-treelet bottom splits; the tree is the same for every thread count.
int orc_bvh_build_threads(const float *tri, uint64_t n, int threads, orc_bvh **out, orc_build_info *info) {
    auto *bvh = new orc_bvh();
    bvh->tri = tri;
    bvh->n = n;
    *out = bvh;
    if (n == 0) return 0;  // hlbvh.cu:392-395: empty aggregate
    if (n >= (1ull << 31)) {
        bvh->error = "too many primitives";
        return -1;
    }
    threads = std::max(1, threads);
    compute_morton(tri, n, bvh->prims, threads);
    canonical_sort(bvh->prims, threads);

    Builder b{bvh};
    for (uint32_t i = 0; i < n;) {  // treelet runs of equal code >> 9 (hlbvh.cu:504-543)
        uint32_t key = bvh->prims[i].code >> kTreeletShift;
        uint32_t j = i;
        B3 bb = b3_empty();
        while (j < n && (bvh->prims[j].code >> kTreeletShift) == key) {
            bb = b3_union(bb, bvh->prims[j].bounds);
            ++j;
        }
        b.treelets.push_back({i, j - i, bb});
        i = j;
    }
    std::vector<uint32_t> ids(b.treelets.size());
    for (uint32_t i = 0; i < ids.size(); ++i) ids[i] = i;
    bvh->nodes.reset(2 * n);
    b.new_nodes(1);
    int spawn = 0;
    while (threads > 1 && (1 << spawn) < 4 * threads) ++spawn;  // up to ~4 subtrees per thread
    b.build_upper(0, ids, spawn);
    if (b.failed) {
        bvh->error = b.why;
        return -2;
    }
    bvh->top_nodes = int(bvh->nodes.size());
    parallel_for(uint64_t(bvh->top_nodes), threads, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) b.build_bottom(int(i));
    });
    if (b.failed) {
        bvh->error = b.why;
        return -2;
    }
    if (info) {
        info->num_treelets = b.treelets.size();
        info->top_nodes = bvh->top_nodes;
        info->total_nodes = bvh->nodes.size();
        uint32_t mx = 0;
        for (auto &t : b.treelets) mx = std::max(mx, t.n);
        info->max_treelet = mx;
    }
    return 0;
}

int orc_bvh_build(const float *tri, uint64_t n, orc_bvh **out, orc_build_info *info) {
    return orc_bvh_build_threads(tri, n, 1, out, info);
}

const char *orc_bvh_error(const orc_bvh *bvh) { return bvh ? bvh->error.c_str() : "null handle"; }

void orc_bvh_free(orc_bvh *bvh) { delete bvh; }

void orc_bvh_set_custom(orc_bvh *bvh, orc_prim_test fn, void *user, const uint8_t *custom) {
    bvh->custom_test = fn;
    bvh->custom_user = user;
    bvh->custom = custom;
}

int orc_bvh_export_canonical(const orc_bvh *bvh, uint8_t *buf, uint64_t *len) {
    std::vector<uint8_t> out;
    std::vector<uint32_t> tmp;
    if (!bvh->nodes.empty()) export_node(bvh, 0, out, tmp);
    if (buf == nullptr) {
        *len = out.size();
        return 0;
    }
    if (*len < out.size()) return -1;
    std::memcpy(buf, out.data(), out.size());
    *len = out.size();
    return 0;
}

// rays: [n][8] floats = o.xyz, t_max, d.xyz, unused. A ray with !(t_max > 0)
// is inactive and reported as a miss.
int orc_bvh_intersect(const orc_bvh *bvh, const float *rays, uint64_t n, int any_hit, int threads,
                      uint32_t *prim_out, float *t_out, float *bary_out, uint64_t *stats_out) {
    std::atomic<int> overflow{0};
    std::atomic<uint64_t> tot_nodes{0}, tot_tris{0};
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        TraceStats st;
        for (uint64_t i = b; i < e; ++i) {
            const float *r = rays + 8 * i;
            V3 o{r[0], r[1], r[2]}, d{r[4], r[5], r[6]};
            float tmax = r[3];
            uint32_t prim = 0xFFFFFFFFu;
            TriHit h{0, 0, 0, 0};
            int rc = (tmax > 0.f) ? trace(bvh, o, d, tmax, any_hit != 0, &prim, &h, &st) : 0;
            if (rc < 0) overflow = 1;
            prim_out[i] = prim;
            if (t_out) t_out[i] = prim == 0xFFFFFFFFu ? 0.f : h.t;
            if (bary_out) {
                bary_out[3 * i] = h.b0;
                bary_out[3 * i + 1] = h.b1;
                bary_out[3 * i + 2] = h.b2;
            }
        }
        tot_nodes += st.nodes;
        tot_tris += st.tris;
    });
    if (stats_out) {
        stats_out[0] = tot_nodes;
        stats_out[1] = tot_tris;
    }
    return overflow ? -1 : 0;
}

// Brute force closest hit over every triangle with the same triangle test.
int orc_brute_intersect(const float *tri, uint64_t ntri, const float *rays, uint64_t n,
                        uint32_t *prim_out, float *t_out) {
    for (uint64_t i = 0; i < n; ++i) {
        const float *r = rays + 8 * i;
        V3 o{r[0], r[1], r[2]}, d{r[4], r[5], r[6]};
        float best = r[3];
        uint32_t prim = 0xFFFFFFFFu;
        for (uint64_t k = 0; k < ntri && best > 0.f; ++k) {
            TriHit h;
            if (intersect_triangle(o, d, best, tri_point(tri, k, 0), tri_point(tri, k, 1),
                                   tri_point(tri, k, 2), &h)) {
                best = h.t;
                prim = uint32_t(k);
            }
        }
        prim_out[i] = prim;
        t_out[i] = prim == 0xFFFFFFFFu ? 0.f : best;
    }
    return 0;
}

// One triangle test, for known-answer checks.
int orc_intersect_triangle(const float *ray, const float *tri9, float t_max, float *out4) {
    TriHit h;
    V3 o{ray[0], ray[1], ray[2]}, d{ray[4], ray[5], ray[6]};
    bool hit = intersect_triangle(o, d, t_max, {tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]},
                                  {tri9[6], tri9[7], tri9[8]}, &h);
    if (hit) {
        out4[0] = h.b0; out4[1] = h.b1; out4[2] = h.b2; out4[3] = h.t;
    }
    return hit ? 1 : 0;
}

// integrators/ambient_occlusion.cu:28-52 with a closest-hit bounce: from each
// primary hit, a cosine-hemisphere direction around the face-forwarded normal,
// u = two PCG32 draws of sequence = ray index, origin offset by spawn_ray.
int orc_ao_bounce(const float *tri, const float *rays, const uint32_t *prim, const float *bary3,
                   uint64_t n, float *rays_out) {
    const float pi = float(std::acos(-1.0));
    for (uint64_t i = 0; i < n; ++i) {
        float *ro = rays_out + 8 * i;
        for (int k = 0; k < 8; ++k) ro[k] = 0.f;
        if (prim[i] == 0xFFFFFFFFu) continue;
        const float *r = rays + 8 * i;
        V3 rd{r[4], r[5], r[6]};
        V3 p0 = tri_point(tri, prim[i], 0), p1 = tri_point(tri, prim[i], 1), p2 = tri_point(tri, prim[i], 2);
        TriHit h{bary3[3 * i], bary3[3 * i + 1], bary3[3 * i + 2], 0};
        HitGeom g = hit_geometry(h, p0, p1, p2);
        V3 nrm = g.n;
        V3 ff = dot(nrm, neg(rd)) < 0.0 ? neg(nrm) : nrm;  // vector3.h:156-158
        Pcg32 rng;
        rng.set_sequence(i);
        float u0 = rng.next_float();
        float u1 = rng.next_float();
        V3 local = sample_cosine_hemisphere(u0, u1);
        float pdf = std::fabs(local.z) / pi;
        if (pdf == 0.0f) continue;
        V3 fx, fy;
        coordinate_system(ff, &fx, &fy);
        V3 wi = add(add(scale(local.x, fx), scale(local.y, fy)), scale(local.z, ff));
        V3 org = offset_ray_origin(g.pi, g.n, wi);
        ro[0] = org.x; ro[1] = org.y; ro[2] = org.z;
        ro[3] = std::numeric_limits<float>::infinity();
        ro[4] = wi.x; ro[5] = wi.y; ro[6] = wi.z;
    }
    return 0;
}

// mode 0: default state/stream; 1: set_sequence(seq) (offset = mix_bits(seq));
// 2: set_sequence(seq, offset).
uint32_t orc_pcg32_sequence(uint64_t seq, uint64_t offset, int mode, int64_t advance, uint32_t *out,
                            uint64_t count) {
    Pcg32 r;
    if (mode == 1) r.set_sequence(seq);
    if (mode == 2) r.set_sequence(seq, offset);
    if (advance) r.advance(advance);
    for (uint64_t i = 0; i < count; ++i) out[i] = r.next_u32();
    return 0;
}

int orc_trace_ray(const orc_bvh *bvh, const float o[3], const float d[3], float t_max, int any_hit,
                  uint32_t *prim, float *t, float *bary3) {
    TriHit h{};
    TraceStats st;
    const int rc = trace(bvh, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), t_max, any_hit != 0, prim, &h, &st);
    if (rc < 0) return -1;
    if (t) *t = rc ? h.t : 0.f;
    if (bary3) {
        bary3[0] = rc ? h.b0 : 0.f;
        bary3[1] = rc ? h.b1 : 0.f;
        bary3[2] = rc ? h.b2 : 0.f;
    }
    return rc;
}

uint64_t orc_murmur64a(const uint8_t *key, uint64_t len, uint64_t seed) {
    return murmur64a(key, len, seed);
}

}  // extern "C"
