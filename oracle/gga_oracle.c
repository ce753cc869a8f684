/*
 * gga_oracle.c -- CPU oracle (test infrastructure only; see gga_oracle.h).
 *
 * Citations are to /Programs/CUDA_src/Kepler_BFS_SSSP/ of
 * vivekkumark/GPUGraphAlgo.  Build: oracle/Makefile (plain gcc, no -ffast-math,
 * SSE float32 arithmetic so the histogram heuristic rounds like the reference).
 */
#include "gga_oracle.h"

#include <stdlib.h>
#include <string.h>

/* graph.h:115-184 (findSerialSSSP / findSerialBFS). */
uint32_t orc_serial_sp(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                       const uint32_t *w, uint32_t src, uint32_t inf,
                       uint32_t *dist, uint32_t *max_dist)
{
    uint32_t sweeps = 0;
    for (uint32_t i = 0; i < n; i++) dist[i] = inf;       /* graph.h:117-118 */
    if (src < n) dist[src] = 0;                            /* graph.h:120 */
    for (uint32_t i = 0; i < n; i++) {                     /* graph.h:122 */
        int flag = 1;
        sweeps++;
        for (uint32_t ver = 0; ver < n; ver++) {           /* graph.h:124 */
            uint32_t tdist = dist[ver];                    /* read once per vertex, :125 */
            if (tdist >= inf) continue;                    /* inf + w never relaxes */
            for (uint32_t ed = row_ptr[ver]; ed < row_ptr[ver + 1]; ed++) {
                uint32_t wt = w ? w[ed] : 1u;              /* :127 / :163 */
                uint32_t alt = tdist + wt;
                if (alt < tdist || alt >= inf) continue;   /* would wrap or reach inf */
                uint32_t d = dst[ed];
                if (alt < dist[d]) {                       /* :128 */
                    flag = 0;
                    dist[d] = alt;                         /* :130 */
                }
            }
        }
        if (flag) break;                                   /* :134-135 */
    }
    if (max_dist) {                                        /* :138-145 */
        uint32_t mx = 0;
        for (uint32_t v = 0; v < n; v++)
            if (dist[v] < inf && dist[v] > mx) mx = dist[v];
        *max_dist = mx;
    }
    return sweeps;
}

/* myutils.h:1-18 (verifysolution). */
uint64_t orc_verify(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                    const uint32_t *w, const uint32_t *dist, uint32_t inf)
{
    uint64_t nerr = 0;
    for (uint32_t u = 0; u < n; u++) {
        uint32_t du = dist[u];
        if (du >= inf) continue;
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
            uint32_t wt = w ? w[e] : 1u;
            if (wt == 0) continue;                         /* `wt > 0 &&`, :12 */
            uint32_t alt = du + wt;
            if (alt < du || alt >= inf) continue;
            if (alt < dist[dst[e]]) nerr++;
        }
    }
    return nerr;
}

uint64_t orc_tight(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                   const uint32_t *w, uint32_t src, const uint32_t *dist, uint32_t inf)
{
    unsigned char *tight = (unsigned char *)calloc(n ? n : 1, 1);
    uint64_t bad = 0;
    for (uint32_t u = 0; u < n; u++) {
        uint32_t du = dist[u];
        if (du >= inf) continue;
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
            uint32_t wt = w ? w[e] : 1u;
            uint32_t alt = du + wt;
            if (alt < du) continue;
            if (alt == dist[dst[e]]) tight[dst[e]] = 1;
        }
    }
    for (uint32_t v = 0; v < n; v++)
        if (v != src && dist[v] < inf && !tight[v]) bad++;
    if (src < n && dist[src] != 0) bad++;
    free(tight);
    return bad;
}

/* graph.h:338-358: strict `<` keeps the first maximum. */
uint32_t orc_maxdeg(uint32_t n, const uint32_t *row_ptr, uint32_t *first_index)
{
    uint32_t best = 0, idx = 0;
    if (n) best = row_ptr[1] - row_ptr[0];
    for (uint32_t j = 1; j < n; j++) {
        uint32_t d = row_ptr[j + 1] - row_ptr[j];
        if (best < d) { best = d; idx = j; }
    }
    if (first_index) *first_index = idx;
    return best;
}

/* graph.h:390: MAX_EDGES_ALLOWED = (maxdeg + NodeSplitLevel - 1) / NodeSplitLevel,
 * evaluated in unsigned arithmetic (maxdeg is unsigned). */
uint32_t orc_mda_level(uint32_t maxdeg, int32_t level)
{
    uint32_t L = (uint32_t)level;
    if (L == 0) return 0;
    return (maxdeg + L - 1u) / L;
}

static uint32_t histo_pick(const uint32_t *histo, int32_t buckets, uint32_t maxdeg)
{
    /* graph.h:375-384 */
    int max = -1, maxindex = -1;
    for (int i = 0; i < buckets; i++) {
        if ((int)histo[i] > max) { max = (int)histo[i]; maxindex = i; }
    }
    maxindex++;
    int maxdeg1 = (int)maxdeg + 1;
    return (uint32_t)((maxdeg1 / buckets) * maxindex);
}

static int bucket_index(uint32_t deg, int maxdeg1, int32_t buckets)
{
    /* graph.h:365: (noutgoing[i] / (float) maxdeg1) * maxHisto, truncated to int.
     * volatile keeps each float32 rounding step. */
    volatile float q = (float)deg / (float)maxdeg1;
    volatile float s = q * (float)buckets;
    return (int)s;
}

/* graph.h:359-385 (getMaxAllowedOutDegree). */
uint32_t orc_mda_histogram(uint32_t n, const uint32_t *deg, int32_t buckets)
{
    if (buckets <= 0) return 0;
    uint32_t maxdeg = 0;
    for (uint32_t i = 0; i < n; i++) if (deg[i] > maxdeg) maxdeg = deg[i];
    uint32_t *histo = (uint32_t *)calloc((size_t)buckets, sizeof(uint32_t));
    int maxdeg1 = (int)maxdeg + 1;
    for (uint32_t i = 0; i < n; i++) histo[bucket_index(deg[i], maxdeg1, buckets)]++;
    uint32_t r = histo_pick(histo, buckets, maxdeg);
    free(histo);
    return r;
}

uint32_t orc_mda_histogram_dd(uint32_t nrows, const uint32_t *deg, const uint32_t *cnt,
                              int32_t buckets)
{
    if (buckets <= 0) return 0;
    uint32_t maxdeg = 0;
    for (uint32_t i = 0; i < nrows; i++) if (cnt[i] && deg[i] > maxdeg) maxdeg = deg[i];
    uint32_t *histo = (uint32_t *)calloc((size_t)buckets, sizeof(uint32_t));
    int maxdeg1 = (int)maxdeg + 1;
    for (uint32_t i = 0; i < nrows; i++) histo[bucket_index(deg[i], maxdeg1, buckets)] += cnt[i];
    uint32_t r = histo_pick(histo, buckets, maxdeg);
    free(histo);
    return r;
}

/* graph.h:399-412. */
uint64_t orc_split_count(uint32_t n, const uint32_t *deg, uint32_t mda)
{
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t nout = deg[i];
        if (nout == 0) { total++; continue; }
        total += (nout + mda - 1u) / mda;
    }
    return total;
}

/* graph.h:415-522 with INT_SPLIT (bfs_sssp.cu:60). */
uint32_t orc_node_split(uint32_t n, const uint32_t *row_ptr, uint32_t mda,
                        uint32_t *chunk_begin, uint32_t *owner,
                        uint32_t *child_count, uint32_t *presum)
{
    uint32_t c = 0;            /* MY_VERTICE_INDEX */
    uint32_t next_begin = 0;   /* psrc of the next node, 0-based */
    if (presum) presum[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t deg = row_ptr[i + 1] - row_ptr[i];
        if (deg == 0) {                                    /* :433-448 */
            if (presum) presum[i + 1] = presum[i];
            if (chunk_begin) chunk_begin[c] = next_begin;
            if (owner) owner[c] = i;
            if (child_count) child_count[c] = 0;
            c++;
            continue;
        }
        uint32_t split = (deg + mda - 1u) / mda;           /* :451 */
        if (presum) presum[i + 1] = presum[i] + split - 1; /* :452-456 */
        if (child_count) child_count[c] = split - 1;
        uint32_t local_max = (deg + split - 1u) / split;   /* :462 */
        uint32_t so_far = 0;
        for (uint32_t k = 0; k < split; k++) {             /* :468-495 */
            uint32_t count;
            if (k == split - 1) count = deg - so_far;      /* last chunk: remainder */
            else { count = local_max; so_far += count; }
            if (chunk_begin) chunk_begin[c] = next_begin;
            if (owner) owner[c] = i;
            if (child_count && k > 0) child_count[c] = 0;
            next_begin += count;
            c++;
        }
    }
    if (chunk_begin) chunk_begin[c] = next_begin;          /* :498 */
    return c;
}

/* The reference's GPU path, sequentially: NodeSplit + init + M_my_drelax sweeps. */
uint32_t orc_split_relax(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                         const uint32_t *w, uint32_t mda, uint32_t src, uint32_t inf,
                         uint32_t *dist)
{
    uint32_t m = row_ptr[n];
    uint64_t C64 = 0;
    {
        uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        for (uint32_t v = 0; v < n; v++) deg[v] = row_ptr[v + 1] - row_ptr[v];
        C64 = orc_split_count(n, deg, mda ? mda : 1);
        free(deg);
    }
    uint32_t C = (uint32_t)C64;
    uint32_t *cb = (uint32_t *)malloc(sizeof(uint32_t) * (C + 1));
    uint32_t *own = (uint32_t *)malloc(sizeof(uint32_t) * (C ? C : 1));
    uint32_t *cc = (uint32_t *)malloc(sizeof(uint32_t) * (C ? C : 1));
    uint32_t *presum = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    orc_node_split(n, row_ptr, mda ? mda : 1, cb, own, cc, presum);
    /* destinations remapped to the new parent ids (graph.h:500-502) */
    uint32_t *ndst = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    for (uint32_t e = 0; e < m; e++) ndst[e] = dst[e] + presum[dst[e]];
    uint32_t *d = (uint32_t *)malloc(sizeof(uint32_t) * (C ? C : 1));
    for (uint32_t i = 0; i < C; i++) d[i] = inf;
    if (src < n) {                                         /* bfs_sssp.cu:302-318 */
        uint32_t p = src + presum[src];
        d[p] = 0;
        for (uint32_t k = 1; k <= cc[p]; k++) d[p + k] = 0;
    }
    uint32_t sweeps = 0;
    int changed = 1;
    while (changed) {                                      /* bfs_sssp.cu:372-468 */
        changed = 0;
        sweeps++;
        for (uint32_t u = 0; u < C; u++) {                 /* cudaKernels.h:285-344 */
            for (uint32_t e = cb[u]; e < cb[u + 1]; e++) {
                uint32_t v = ndst[e];
                uint32_t wt = w ? w[e] : 1u;
                uint32_t du = d[u];
                if (du >= inf) break;
                uint32_t alt = du + wt;
                if (alt < du || alt >= inf) continue;
                if (alt < d[v]) {
                    if (alt < d[v]) d[v] = alt;            /* atomicMin(&dist[v], alt) */
                    for (uint32_t k = 1; k <= cc[v]; k++)  /* atomicMin on each child */
                        if (alt < d[v + k]) d[v + k] = alt;
                    changed = 1;
                }
            }
        }
    }
    for (uint32_t v = 0; v < n; v++) dist[v] = d[v + presum[v]];   /* parents only */
    free(cb); free(own); free(cc); free(presum); free(ndst); free(d);
    return sweeps;
}

/* ---- textbook routines ---- */
typedef struct { uint32_t key, v; } heap_item;

static void heap_push(heap_item *h, uint64_t *sz, heap_item it)
{
    uint64_t i = (*sz)++;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (h[p].key <= it.key) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = it;
}

static heap_item heap_pop(heap_item *h, uint64_t *sz)
{
    heap_item top = h[0];
    heap_item last = h[--(*sz)];
    uint64_t i = 0, n = *sz;
    for (;;) {
        uint64_t l = 2 * i + 1;
        if (l >= n) break;
        uint64_t r = l + 1, c = (r < n && h[r].key < h[l].key) ? r : l;
        if (h[c].key >= last.key) break;
        h[i] = h[c];
        i = c;
    }
    if (n) h[i] = last;
    return top;
}

void orc_dijkstra(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                  const uint32_t *w, uint32_t src, uint32_t inf, uint32_t *dist)
{
    uint32_t m = row_ptr[n];
    for (uint32_t i = 0; i < n; i++) dist[i] = inf;
    if (src >= n) return;
    heap_item *h = (heap_item *)malloc(sizeof(heap_item) * ((uint64_t)m + 2));
    uint64_t sz = 0;
    dist[src] = 0;
    heap_item s0 = {0, src};
    heap_push(h, &sz, s0);
    while (sz) {
        heap_item it = heap_pop(h, &sz);
        if (it.key != dist[it.v]) continue;
        for (uint32_t e = row_ptr[it.v]; e < row_ptr[it.v + 1]; e++) {
            uint32_t wt = w ? w[e] : 1u;
            uint32_t alt = it.key + wt;
            if (alt < it.key || alt >= inf) continue;
            uint32_t v = dst[e];
            if (alt < dist[v]) {
                dist[v] = alt;
                heap_item ni = {alt, v};
                heap_push(h, &sz, ni);
            }
        }
    }
    free(h);
}

void orc_bfs_queue(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                   uint32_t src, uint32_t inf, uint32_t *dist)
{
    for (uint32_t i = 0; i < n; i++) dist[i] = inf;
    if (src >= n) return;
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t head = 0, tail = 0;
    dist[src] = 0;
    q[tail++] = src;
    while (head < tail) {
        uint32_t u = q[head++];
        uint32_t nd = dist[u] + 1;
        if (nd >= inf || nd == 0) continue;
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
            uint32_t v = dst[e];
            if (dist[v] == inf) { dist[v] = nd; q[tail++] = v; }
        }
    }
    free(q);
}

uint64_t orc_m_reach(uint32_t n, const uint32_t *row_ptr, const uint32_t *dist, uint32_t inf)
{
    uint64_t s = 0;
    for (uint32_t v = 0; v < n; v++)
        if (dist[v] < inf) s += row_ptr[v + 1] - row_ptr[v];
    return s;
}
