/*
 * onco_oracle_int.c — CPU oracle, integer half: Morton keys, stable LSD radix
 * sort, Blelloch exclusive scan, bin start/end, Philox4x32-10.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for liboncohip.so; it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and never by the product path.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository). How each one is pinned independently of this
 * restatement is listed beside it ("pinned by").
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Morton key of voxel (x, y, z), 10 bits per axis.
 * Restates framework/src/oncolytica/core/runtime/_geometry.py:14-42
 * (interleave_9bit of c/2, shifted left by 3, or-ed with the sub-voxel index
 * (x&1)|(y&1)<<1|(z&1)<<2), which is the textbook 3-D Morton code of the
 * 10-bit coordinates; the WGSL twin is gpu/compiler/template.wgsl:100-132.
 * Pinned by: tests compare against a bit-by-bit interleave over all 64^3
 * coordinates and against the test_gpu.py-style bucket membership golden. */
static uint32_t spread9(uint32_t v) {
    v &= 0x1FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

uint32_t oracle_morton30(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t j = spread9(x >> 1) | (spread9(y >> 1) << 1) | (spread9(z >> 1) << 2);
    uint32_t i = (x & 1u) | ((y & 1u) << 1) | ((z & 1u) << 2);
    return (j << 3) | i;
}

/* 32-bit variant: bits 0..29 as above, x bit 10 -> bit 30, y bit 10 -> bit 31.
 * No reference counterpart (the reference stops at 30 bits); pinned by the
 * bit-loop test over coordinates up to 2047. */
uint32_t oracle_morton32(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t k = oracle_morton30(x & 0x3FFu, y & 0x3FFu, z & 0x3FFu);
    k |= ((x >> 10) & 1u) << 30;
    k |= ((y >> 10) & 1u) << 31;
    return k;
}

uint32_t oracle_morton(uint32_t x, uint32_t y, uint32_t z, int fmt) {
    return fmt == 32 ? oracle_morton32(x, y, z) : oracle_morton30(x, y, z);
}

/* Number of significant key bits for a grid: 3 * ceil(log2(max_dim)), and
 * passes = ceil(bits / 8). Keys never exceed Morton(dims-1), so sorting those
 * bits yields the same permutation as sorting all 32. */
int oracle_key_bits(const int32_t dims[3]) {
    int32_t m = dims[0];
    if (dims[1] > m) m = dims[1];
    if (dims[2] > m) m = dims[2];
    int b = 0;
    while ((1 << b) < m) b++;
    if (b == 0) b = 1;
    return 3 * b;
}

/* Stable LSD radix sort of (key, val) by the low key_bits bits, 8-bit digits,
 * one counting-sort pass per digit. Replaces the reference's unordered atomic
 * counting sort (gpu/compiler/template.wgsl:165-186 count/offset, 335-343
 * scatter) with the stable order the deterministic spec requires: equal keys
 * keep input order. Pinned by: numpy argsort(kind="stable"). */
void oracle_radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits) {
    if (n == 0) return;
    uint32_t* k2 = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* v2 = (uint32_t*)malloc((size_t)n * 4);
    int passes = (key_bits + 7) / 8;
    for (int p = 0; p < passes; p++) {
        uint32_t cnt[256];
        memset(cnt, 0, sizeof cnt);
        int sh = 8 * p;
        for (uint32_t i = 0; i < n; i++) cnt[(keys[i] >> sh) & 255u]++;
        uint32_t run = 0;
        for (int d = 0; d < 256; d++) {
            uint32_t c = cnt[d];
            cnt[d] = run;
            run += c;
        }
        for (uint32_t i = 0; i < n; i++) {
            uint32_t d = (keys[i] >> sh) & 255u;
            uint32_t o = cnt[d]++;
            k2[o] = keys[i];
            v2[o] = vals[i];
        }
        memcpy(keys, k2, (size_t)n * 4);
        memcpy(vals, v2, (size_t)n * 4);
    }
    free(k2);
    free(v2);
}

/* Exclusive prefix sum by the work-efficient up-sweep / down-sweep of
 * Blelloch (1990), the algorithm of gpu/compiler/template.wgsl:192-222,
 * applied to the whole array padded to a power of two (the reference's three
 * 256-wide levels, :228-329, compose to the same result). Returns the total.
 * Pinned by: numpy cumsum and the invariants out[0]=0, out[i+1]-out[i]=in[i]. */
uint32_t oracle_exscan_u32(const uint32_t* in, uint32_t* out, uint32_t n) {
    if (n == 0) return 0;
    size_t m = 1;
    while (m < n) m <<= 1;
    uint32_t* b = (uint32_t*)calloc(m, 4);
    memcpy(b, in, (size_t)n * 4);
    for (size_t s = 1; s < m; s <<= 1)
        for (size_t i = 2 * s - 1; i < m; i += 2 * s) b[i] += b[i - s];
    uint32_t total = b[m - 1];
    b[m - 1] = 0;
    for (size_t s = m >> 1; s >= 1; s >>= 1) {
        for (size_t i = 2 * s - 1; i < m; i += 2 * s) {
            uint32_t t = b[i - s];
            b[i - s] = b[i];
            b[i] += t;
        }
        if (s == 1) break;
    }
    memcpy(out, b, (size_t)n * 4);
    free(b);
    return total;
}

/* Bin start/end over sorted keys: for every key k present, start[k] is the
 * first index holding k and end[k] one past the last. Entries of absent keys
 * are not written. Replaces VoxelTable{count,startIndex} (template.wgsl:42-45,
 * 228-329), whose startIndex[k] = #keys < k and count[k] = #keys == k agree
 * with (start, end - start) on every occupied key.
 * Pinned by: numpy searchsorted left/right on the occupied keys. */
void oracle_bins(const uint32_t* keys, uint32_t n, uint32_t* start, uint32_t* end) {
    for (uint32_t i = 0; i < n; i++) {
        if (i == 0 || keys[i - 1] != keys[i]) start[keys[i]] = i;
        if (i + 1 == n || keys[i + 1] != keys[i]) end[keys[i]] = i + 1;
    }
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"): ten rounds of the 4x32 bijection with multipliers
 * 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85.
 * No reference counterpart: the reference draws from CPython's MT19937
 * (core/utils/_types.py:703-712) or per-entity xorshift32
 * (gpu/compiler/template.wgsl:75-98); the spec replaces both by Philox keyed
 * (seed, cell id, step, call block).
 * Pinned by: Random123's published known-answer vectors and rocRAND's
 * philox4x32_10_engine::ten_rounds (tests/golden/philox_kat.json). */
void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

void oracle_philox_batch(const uint32_t* ctr, uint32_t n, uint64_t key, uint32_t* out) {
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    for (uint32_t i = 0; i < n; i++) oracle_philox4x32_10(ctr + 4 * (size_t)i, k, out + 4 * (size_t)i);
}

void oracle_morton_batch(const uint32_t* x, const uint32_t* y, const uint32_t* z, uint32_t n, int fmt,
                         uint32_t* out) {
    for (uint32_t i = 0; i < n; i++) out[i] = oracle_morton(x[i], y[i], z[i], fmt);
}

/* Threads the per-cell and per-voxel loops of the oracle step use (OpenMP;
 * results do not depend on it: every parallel loop writes disjoint outputs). */
#ifdef _OPENMP
#include <omp.h>
int oracle_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
#else
int oracle_threads(void) { return 1; }
void oracle_set_threads(int n) { (void)n; }
#endif
