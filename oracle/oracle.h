/* TEST INFRASTRUCTURE ONLY — the CPU oracle's C interface.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle.so, and only as the checker or the timed CPU baseline; the product
 * library (libpbrtm.so) never links or calls it.
 *
 * Parity status per function (how each is pinned — tests/test_oracle_*.py):
 *   orc_pcg32_sequence   pinned: PCG32 reference known-answer outputs, advance == stepping
 *   orc_murmur64a        pinned: MurmurHash64A known-answer vectors, length/tail cases
 *   orc_morton_codes     pinned: brute-force bit interleave + closed-form cell index
 *   orc_morton_sort      pinned: permutation, sortedness, tie order (numpy lexsort)
 *   orc_bvh_build        pinned: brute-force recursive SAH/midpoint restatement on tiny
 *                        inputs (Python), set-tree invariants on large ones
 *   orc_intersect_triangle pinned: FP64 Moller-Trumbore within 1e-6, watertight edge rays
 *   orc_bvh_intersect    pinned: equals orc_brute_intersect (prim and t) on seeded rays
 *   orc_ao_bounce        pinned: unit direction, hemisphere side, cosine distribution
 *                        (chi-square), origin within the error box of the hit point
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_bvh orc_bvh;

typedef struct orc_build_info {
    uint64_t num_treelets;
    uint64_t top_nodes;
    uint64_t total_nodes;
    uint64_t max_treelet;
} orc_build_info;

int orc_morton_codes(const float *tri, uint64_t n, uint32_t *codes_by_input);
int orc_morton_sort(const float *tri, uint64_t n, uint32_t *codes_out, uint32_t *order_out);
int orc_bvh_build(const float *tri, uint64_t n, orc_bvh **out, orc_build_info *info);
/* the same tree, built with that many host threads (the CPU baseline's build) */
int orc_bvh_build_threads(const float *tri, uint64_t n, int threads, orc_bvh **out, orc_build_info *info);
const char *orc_bvh_error(const orc_bvh *bvh);
void orc_bvh_free(orc_bvh *bvh);
int orc_bvh_export_canonical(const orc_bvh *bvh, uint8_t *buf, uint64_t *len);
int orc_bvh_intersect(const orc_bvh *bvh, const float *rays, uint64_t n, int any_hit, int threads,
                      uint32_t *prim_out, float *t_out, float *bary_out, uint64_t *stats_out);
int orc_brute_intersect(const float *tri, uint64_t ntri, const float *rays, uint64_t n,
                        uint32_t *prim_out, float *t_out);
int orc_intersect_triangle(const float *ray, const float *tri9, float t_max, float *out4);
int orc_ao_bounce(const float *tri, const float *rays, const uint32_t *prim, const float *bary3,
                  uint64_t n, float *rays_out);
uint32_t orc_pcg32_sequence(uint64_t seq, uint64_t offset, int mode, int64_t advance, uint32_t *out,
                            uint64_t count);
uint64_t orc_murmur64a(const uint8_t *key, uint64_t len, uint64_t seed);
/* Primitives that are not triangles (analytic shapes, instances): custom[i] != 0
 * routes primitive i's test to fn, which returns 1 on a hit with t < or = t_max
 * as the shape defines it, writing t and three words for the hit record. */
typedef int (*orc_prim_test)(void *user, uint32_t prim, const float o[3], const float d[3], float t_max, float *t,
                             float b3[3]);
void orc_bvh_set_custom(orc_bvh *bvh, orc_prim_test fn, void *user, const uint8_t *custom);
/* One ray through the oracle tree (closest or any hit); 1 = hit, 0 = miss, -1 = stack overflow. */
int orc_trace_ray(const orc_bvh *bvh, const float o[3], const float d[3], float t_max, int any_hit,
                  uint32_t *prim, float *t, float *bary3);

#ifdef __cplusplus
}
#endif
