/* Exhaustive check of the slab-test widening used by the HIP traversal.
 *
 * bounds3.h:159-180 widens a far slab distance t (FP32) by the double factor
 * 1.0 + 2.0 * gamma(3) and stores the product back into FP32:
 * float(double(t) * (1 + d)), d = 2 * gamma(3) (an FP32 number). The kernels
 * compute fmaf(t, d, t) instead, which rounds the exact product t * (1 + d)
 * once. The two can differ only where the FP64 rounding lands on an FP32
 * rounding tie; this program visits all 2^32 FP32 bit patterns and counts the
 * patterns where they differ (NaN results compare equal to each other).
 * Prints "mismatches <count>" and exits 0 when the count is zero. */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define THREADS 8

static float gamma_n(int n) {
    const float eps = FLT_EPSILON * 0.5f; /* 2^-24 */
    return ((float)n * eps) / (1.0f - (float)n * eps);
}

static double factor;
static float delta;
static uint64_t bad[THREADS];

static void *scan(void *arg) {
    const uint64_t k = (uint64_t)(intptr_t)arg;
    const uint64_t lo = k << 29, hi = (k + 1) << 29; /* 2^32 / THREADS */
    uint64_t nb = 0;
    for (uint64_t u = lo; u < hi; ++u) {
        const uint32_t b = (uint32_t)u;
        float t;
        memcpy(&t, &b, 4);
        const float ref = (float)((double)t * factor);
        const float fused = fmaf(t, delta, t);
        uint32_t br, bf;
        memcpy(&br, &ref, 4);
        memcpy(&bf, &fused, 4);
        if (br != bf && !(isnan(ref) && isnan(fused))) ++nb;
    }
    bad[k] = nb;
    return NULL;
}

int main(void) {
    factor = 1.0 + 2.0 * (double)gamma_n(3);
    delta = 2.0f * gamma_n(3);
    if ((double)delta != factor - 1.0) {
        printf("factor - 1 is not the FP32 number 2 * gamma(3)\n");
        return 2;
    }
    pthread_t th[THREADS];
    for (int k = 0; k < THREADS; ++k) pthread_create(&th[k], NULL, scan, (void *)(intptr_t)k);
    uint64_t total = 0;
    for (int k = 0; k < THREADS; ++k) {
        pthread_join(th[k], NULL);
        total += bad[k];
    }
    printf("mismatches %llu\n", (unsigned long long)total);
    return total == 0 ? 0 : 1;
}
