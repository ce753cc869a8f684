/*
 * gen.c -- seeded synthetic graph generators (the reference used GTgraph's
 * R-MAT externally and ships no generator; rmat20.gr is not in the repo).
 *
 * Randomness is counter-based (splitmix64 of (seed, stream, index)), so the
 * output is a pure function of the arguments whatever the thread count.
 *
 * R-MAT: arc i draws `scale` quadrant choices from 32-bit uniforms u compared
 * with floor(a * 2^32), floor((a + b) * 2^32), floor((a + b + c) * 2^32):
 * quadrant a keeps (src, dst) bits (0, 0), b (0, 1), c (1, 0), d (1, 1).  No
 * vertex permutation (hubs sit at low ids, as in GTgraph output).  Arcs are
 * sorted by (src, dst); with dedup, duplicates and self-loops are removed.
 * Weight of arc (u, v) = wlo + floor(h * (whi - wlo + 1) / 2^32) with h the
 * high half of splitmix64(seed' ^ (u << 32 | v)), so a duplicate arc has the
 * same weight as its twin.
 *
 * Grid: vertex (r, c) has id r * cols + c; 4-neighbour undirected edges as
 * two arcs; k = round(shortcut_frac * undirected grid edges) shortcuts drawn
 * as uniform pairs {u, v}, rejecting u == v, grid neighbours and repeats;
 * each added in both directions.  One weight per undirected edge from the
 * hash of (min(u, v), max(u, v)).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gga_host.h"

/* ----------------------------------------------------------- radix sort */
int gga_radix_sort_u64(uint64_t *keys, uint64_t count, unsigned bits)
{
    if (count < 2) return GGA_OK;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * count);
    if (!tmp) return GGA_ENOMEM;
    static const unsigned DIG = 16;
    uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) << DIG);
    if (!cnt) {
        free(tmp);
        return GGA_ENOMEM;
    }
    uint64_t *a = keys, *b = tmp;
    unsigned passes = (bits + DIG - 1) / DIG;
    for (unsigned p = 0; p < passes; p++) {
        unsigned sh = p * DIG;
        memset(cnt, 0, sizeof(uint64_t) << DIG);
        for (uint64_t i = 0; i < count; i++) cnt[(a[i] >> sh) & 0xFFFF]++;
        uint64_t s = 0;
        for (uint64_t d = 0; d < (1u << DIG); d++) {
            uint64_t c = cnt[d];
            cnt[d] = s;
            s += c;
        }
        for (uint64_t i = 0; i < count; i++) b[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
        uint64_t *t = a;
        a = b;
        b = t;
    }
    if (a != keys) memcpy(keys, a, sizeof(uint64_t) * count);
    free(cnt);
    free(tmp);
    return GGA_OK;
}

int gga_csr_from_sorted_keys(uint32_t n, const uint64_t *keys, uint64_t count, gga_csr *out)
{
    if (count > UINT32_MAX) return GGA_ERANGE;
    memset(out, 0, sizeof *out);
    out->n = n;
    out->m = (uint32_t)count;
    out->row_ptr = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    out->dst = (uint32_t *)malloc(sizeof(uint32_t) * (count ? count : 1));
    out->w = (uint32_t *)malloc(sizeof(uint32_t) * (count ? count : 1));
    if (!out->row_ptr || !out->dst || !out->w) {
        gga_csr_free(out);
        return GGA_ENOMEM;
    }
    for (uint64_t i = 0; i < count; i++) {
        uint32_t s = (uint32_t)(keys[i] >> 32);
        out->row_ptr[s + 1]++;
        out->dst[i] = (uint32_t)keys[i];
    }
    for (uint32_t v = 0; v < n; v++) out->row_ptr[v + 1] += out->row_ptr[v];
    out->has_weights = 1;
    return GGA_OK;
}

static unsigned hw_threads(void)
{
    long t = sysconf(_SC_NPROCESSORS_ONLN);
    if (t < 1) t = 1;
    if (t > 16) t = 16;
    return (unsigned)t;
}

static inline uint32_t uniform_range(uint64_t h, uint32_t lo, uint32_t hi)
{
    uint64_t span = (uint64_t)hi - lo + 1u;
    return lo + (uint32_t)(((h >> 32) * span) >> 32);
}

/* --------------------------------------------------------------- R-MAT */
typedef struct {
    uint64_t *keys;
    uint64_t begin, end;
    uint32_t scale;
    uint64_t seed;
    uint64_t ta, tab, tabc;
} rmat_job;

static void *rmat_worker(void *arg)
{
    rmat_job *j = (rmat_job *)arg;
    for (uint64_t i = j->begin; i < j->end; i++) {
        uint64_t base = gga_splitmix64(j->seed ^ gga_splitmix64(i));
        uint64_t src = 0, dst = 0, r = 0;
        for (uint32_t lv = 0; lv < j->scale; lv++) {
            if ((lv & 1u) == 0) r = gga_splitmix64(base + lv);
            uint64_t u = (lv & 1u) ? (r >> 32) : (r & 0xFFFFFFFFull);
            unsigned sb, db;
            if (u < j->ta) { sb = 0; db = 0; }
            else if (u < j->tab) { sb = 0; db = 1; }
            else if (u < j->tabc) { sb = 1; db = 0; }
            else { sb = 1; db = 1; }
            src = (src << 1) | sb;
            dst = (dst << 1) | db;
        }
        j->keys[i] = (src << 32) | dst;
    }
    return NULL;
}

typedef struct {
    const uint64_t *keys;
    uint32_t *w;
    uint64_t begin, end;
    uint64_t wseed;
    uint32_t wlo, whi;
    int symmetric;
} weight_job;

static void *weight_worker(void *arg)
{
    weight_job *j = (weight_job *)arg;
    for (uint64_t i = j->begin; i < j->end; i++) {
        uint64_t k = j->keys[i];
        if (j->symmetric) {
            uint64_t u = k >> 32, v = k & 0xFFFFFFFFull;
            k = u < v ? (u << 32 | v) : (v << 32 | u);
        }
        j->w[i] = uniform_range(gga_splitmix64(j->wseed ^ k), j->wlo, j->whi);
    }
    return NULL;
}

static void fill_weights(const uint64_t *keys, uint32_t *w, uint64_t m, uint64_t wseed, uint32_t wlo,
                         uint32_t whi, int symmetric)
{
    unsigned nt = hw_threads();
    pthread_t th[16];
    weight_job jobs[16];
    for (unsigned t = 0; t < nt; t++) {
        jobs[t].keys = keys;
        jobs[t].w = w;
        jobs[t].begin = m * t / nt;
        jobs[t].end = m * (t + 1) / nt;
        jobs[t].wseed = wseed;
        jobs[t].wlo = wlo;
        jobs[t].whi = whi;
        jobs[t].symmetric = symmetric;
        pthread_create(&th[t], NULL, weight_worker, &jobs[t]);
    }
    for (unsigned t = 0; t < nt; t++) pthread_join(th[t], NULL);
}

int gga_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                 uint32_t wlo, uint32_t whi, int dedup, gga_csr *out)
{
    if (!out || scale == 0 || scale > 31 || wlo > whi || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return GGA_EINVAL;
    memset(out, 0, sizeof *out);
    const uint32_t n = 1u << scale;
    const uint64_t m0 = (uint64_t)edge_factor << scale;
    if (m0 > UINT32_MAX) return GGA_ERANGE;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (m0 ? m0 : 1));
    if (!keys) return GGA_ENOMEM;
    const double two32 = 4294967296.0;
    rmat_job jobs[16];
    pthread_t th[16];
    unsigned nt = hw_threads();
    for (unsigned t = 0; t < nt; t++) {
        jobs[t].keys = keys;
        jobs[t].begin = m0 * t / nt;
        jobs[t].end = m0 * (t + 1) / nt;
        jobs[t].scale = scale;
        jobs[t].seed = gga_splitmix64(seed * 0x100000001B3ull + 0x52A7u);
        jobs[t].ta = (uint64_t)floor(a * two32);
        jobs[t].tab = (uint64_t)floor((a + b) * two32);
        jobs[t].tabc = (uint64_t)floor((a + b + c) * two32);
        pthread_create(&th[t], NULL, rmat_worker, &jobs[t]);
    }
    for (unsigned t = 0; t < nt; t++) pthread_join(th[t], NULL);
    int rc = gga_radix_sort_u64(keys, m0, 32 + scale);
    if (rc) {
        free(keys);
        return rc;
    }
    uint64_t m = m0;
    if (dedup) {
        uint64_t o = 0;
        for (uint64_t i = 0; i < m0; i++) {
            uint64_t k = keys[i];
            if ((k >> 32) == (k & 0xFFFFFFFFull)) continue;   /* self-loop */
            if (o && keys[o - 1] == k) continue;              /* duplicate */
            keys[o++] = k;
        }
        m = o;
    }
    rc = gga_csr_from_sorted_keys(n, keys, m, out);
    if (!rc) fill_weights(keys, out->w, m, gga_splitmix64(seed ^ 0x57E16A7ull), wlo, whi, 0);
    free(keys);
    return rc;
}

/* ---------------------------------------------------------------- grid */
static int is_grid_neighbour(uint64_t u, uint64_t v, uint32_t cols)
{
    if (u > v) {
        uint64_t t = u;
        u = v;
        v = t;
    }
    if (v - u == cols) return 1;
    if (v - u == 1 && (v % cols) != 0) return 1;
    return 0;
}

/* open-addressing set of undirected pairs (key = min << 32 | max) */
typedef struct {
    uint64_t *slot;
    uint64_t mask;
} pairset;

static int pairset_insert(pairset *s, uint64_t key)
{
    uint64_t h = gga_splitmix64(key) & s->mask;
    for (;;) {
        if (s->slot[h] == UINT64_MAX) {
            s->slot[h] = key;
            return 1;
        }
        if (s->slot[h] == key) return 0;
        h = (h + 1) & s->mask;
    }
}

int gga_gen_grid(uint32_t rows, uint32_t cols, double shortcut_frac, uint64_t seed, uint32_t wlo,
                 uint32_t whi, gga_csr *out)
{
    if (!out || rows == 0 || cols == 0 || wlo > whi || shortcut_frac < 0) return GGA_EINVAL;
    memset(out, 0, sizeof *out);
    const uint64_t n64 = (uint64_t)rows * cols;
    if (n64 >= UINT32_MAX) return GGA_ERANGE;
    const uint32_t n = (uint32_t)n64;
    const uint64_t g_edges = (uint64_t)rows * (cols - 1) + (uint64_t)(rows - 1) * cols;
    const uint64_t k = (uint64_t)llround(shortcut_frac * (double)g_edges);
    if (n < 3 && k) return GGA_EINVAL;
    const uint64_t m = 2 * (g_edges + k);
    if (m > UINT32_MAX) return GGA_ERANGE;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (m ? m : 1));
    if (!keys) return GGA_ENOMEM;
    uint64_t o = 0;
    for (uint32_t r = 0; r < rows; r++) {
        for (uint32_t c = 0; c < cols; c++) {
            uint64_t u = (uint64_t)r * cols + c;
            if (c + 1 < cols) {
                keys[o++] = u << 32 | (u + 1);
                keys[o++] = (u + 1) << 32 | u;
            }
            if (r + 1 < rows) {
                keys[o++] = u << 32 | (u + cols);
                keys[o++] = (u + cols) << 32 | u;
            }
        }
    }
    if (k) {
        pairset ps;
        uint64_t cap = 1;
        while (cap < 2 * k + 16) cap <<= 1;
        ps.mask = cap - 1;
        ps.slot = (uint64_t *)malloc(sizeof(uint64_t) * cap);
        if (!ps.slot) {
            free(keys);
            return GGA_ENOMEM;
        }
        memset(ps.slot, 0xFF, sizeof(uint64_t) * cap);
        uint64_t sseed = gga_splitmix64(seed ^ 0x5C0A7C07ull), draw = 0, added = 0;
        while (added < k) {
            uint64_t r = gga_splitmix64(sseed + draw++);
            uint64_t u = ((r & 0xFFFFFFFFull) * n) >> 32, v = ((r >> 32) * n) >> 32;
            if (u == v || is_grid_neighbour(u, v, cols)) continue;
            uint64_t key = u < v ? (u << 32 | v) : (v << 32 | u);
            if (!pairset_insert(&ps, key)) continue;
            keys[o++] = u << 32 | v;
            keys[o++] = v << 32 | u;
            added++;
        }
        free(ps.slot);
    }
    unsigned bits = 32;
    while (bits < 64 && (1ull << (bits - 32)) < n64) bits++;
    int rc = gga_radix_sort_u64(keys, m, bits);
    if (!rc) rc = gga_csr_from_sorted_keys(n, keys, m, out);
    if (!rc) fill_weights(keys, out->w, m, gga_splitmix64(seed ^ 0x6121D7ull), wlo, whi, 1);
    free(keys);
    return rc;
}
