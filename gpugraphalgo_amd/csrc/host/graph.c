/*
 * graph.c -- device-resident graphs, the split driver and the traversal
 * drivers of the C-ABI (include/gga.h).
 *
 * Replaces, from Programs/CUDA_src/Kepler_BFS_SSSP/:
 *   Graph::allocOnDevice / cudaCopy / CudaTest   graph.h:523-553, 683-699
 *   the split driver + MDA selection             bfs_sssp.cu:224-253, graph.h:338-522
 *   distance init, fixpoint loop, timing         bfs_sssp.cu:302-321, 345-473
 * The host never renumbers the graph: node splitting is the chunk table the
 * split kernels build on the device (kernels/split.hip).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../gga_kernels.h"
#include "gga_host.h"

#ifndef GGA_MESH_BMIN
#define GGA_MESH_BMIN 4u    /* chunks per wave batch on mesh-like graphs */
#endif
static int mesh_like(const gga_graph *g);

#define GGA_VERSION "gga 0.1.0 (gfx950)"

struct gga_graph {
    int device;
    hipStream_t stream;
    uint32_t n, m;
    uint32_t *d_row, *d_col, *d_w;
    uint32_t maxdeg, maxdeg_idx;
    /* split */
    int split;
    uint32_t C, mda;              /* C: chunks of the table the kernels walk (d_cstart/d_cbeg/d_cown) */
    uint32_t *d_cstart, *d_cbeg, *d_cown;
    uint32_t Cref;                /* chunks of the reference's NodeSplit (ns_nnode) */
    uint32_t *d_rcbeg, *d_rcown;  /* its descriptors when they differ from the walked table, else NULL */
    gga_split_stats sst;
    /* query workspace */
    uint32_t *d_dist, *d_mark, *d_q0, *d_q1, *d_ctl, *d_tiles;
    uint8_t *d_flag;
    uint64_t *d_cnt;
    uint32_t qcap;
    uint32_t *h_pin;          /* pinned: GGA_CTL_WORDS words */
    uint32_t epoch;           /* step number of the next query's first step (mark[] stamps are below it) */
    int mark_dirty;           /* mark[] holds stamps of unknown step numbers: refill before the next query */
    int have_result;
    uint32_t wl_grid[3][2][2];/* [workgroup size 128 / 256 / 512][bfs][count] */
    uint32_t p_grid[2][2];
    uint32_t split_gen;       /* bumped by every split; engines built on an older table are stale */
    gga_ms *ms_cache;         /* engine of gga_multi_sp */
    uint32_t tuned_grid, tuned_dense, tuned_layout;   /* gga_tune results (0: defaults) */
    uint32_t tuned_block;        /* worklist workgroup size from gga_tune (0: 256) */
    uint32_t tuned_delta;
    int32_t tuned_mode;          /* 0: none; else GGA_MODE_* + 1 */
    uint32_t last_steps[2];      /* [bfs] relax steps of the last query (AUTO's evidence) */
    uint32_t wmean;              /* mean edge weight, rounded up (>= 1) */
    uint32_t *d_last, *d_tmin;   /* near-far state (allocated on first use) */
    void *msbfs;              /* transposed graph + state of gga_multi_bfs (msbfs.c) */
    hipEvent_t ev[4];
    hipEvent_t *kev;          /* per-launch timing events */
    uint32_t kev_cap;
    uint64_t dev_bytes;
};

const char *gga_version(void) { return GGA_VERSION; }

static pthread_mutex_t coop_mu = PTHREAD_MUTEX_INITIALIZER;
void gga_coop_lock(void) { pthread_mutex_lock(&coop_mu); }
void gga_coop_unlock(void) { pthread_mutex_unlock(&coop_mu); }

int gga_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            gga_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
            rc = (e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation) ? GGA_ENOMEM \
                                                                                : GGA_EHIP; \
            goto fail;                                                                     \
        }                                                                                  \
    } while (0)

#define K_TRY(expr)                                                                        \
    do {                                                                                   \
        int e_ = (expr);                                                                   \
        if (e_ != 0) {                                                                     \
            gga_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,                       \
                          hipGetErrorString((hipError_t)e_));                              \
            rc = GGA_EHIP;                                                                 \
            goto fail;                                                                     \
        }                                                                                  \
    } while (0)

static int dmalloc(gga_graph *g, void **p, uint64_t bytes)
{
    *p = NULL;
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        gga_set_error("hipMalloc %llu bytes: %s", (unsigned long long)bytes, hipGetErrorString(e));
        *p = NULL;
        return GGA_ENOMEM;
    }
    g->dev_bytes += bytes;
    return GGA_OK;
}

static void dfree(gga_graph *g, void **p, uint64_t bytes)
{
    if (*p) {
        hipFree(*p);
        if (bytes == 0) bytes = 4;
        g->dev_bytes -= bytes;
    }
    *p = NULL;
}

static gga_dev_graph dev_graph(const gga_graph *g)
{
    gga_dev_graph d;
    d.row = g->d_row;
    d.col = g->d_col;
    d.w = g->d_w;
    d.cstart = g->d_cstart;
    d.cbeg = g->d_cbeg;
    d.cown = g->d_cown;
    d.n = g->n;
    d.m = g->m;
    d.C = g->C;
    d.mda = g->mda;
    return d;
}

static gga_dev_work dev_work(const gga_graph *g, uint32_t inf)
{
    gga_dev_work w;
    w.dist = g->d_dist;
    w.mark = g->d_mark;
    w.flag = g->d_flag;
    w.q0 = g->d_q0;
    w.q1 = g->d_q1;
    w.ctl = g->d_ctl;
    w.cnt = g->d_cnt;
    w.tiles = g->d_tiles;
    w.inf = inf;
    w.qcap = g->qcap;
    w.dense_min = 0xFFFFFFFFu;
    w.last = NULL;
    w.tmin = NULL;
    w.delta = 0;
    w.bmin = 1;
    return w;
}

uint32_t gga_graph_n(const gga_graph *g) { return g ? g->n : 0; }
uint32_t gga_graph_m(const gga_graph *g) { return g ? g->m : 0; }
uint64_t gga_graph_device_bytes(const gga_graph *g) { return g ? g->dev_bytes : 0; }

static void free_split(gga_graph *g)
{
    dfree(g, (void **)&g->d_cstart, 4ull * ((uint64_t)g->n + 1));
    dfree(g, (void **)&g->d_cbeg, 4ull * ((uint64_t)g->C + 1));
    dfree(g, (void **)&g->d_cown, 4ull * g->C);
    dfree(g, (void **)&g->d_rcbeg, 4ull * ((uint64_t)g->Cref + 1));
    dfree(g, (void **)&g->d_rcown, 4ull * g->Cref);
    g->split = 0;
}

static void free_queues(gga_graph *g)
{
    dfree(g, (void **)&g->d_q0, 4ull * g->qcap);
    dfree(g, (void **)&g->d_q1, 4ull * g->qcap);
    g->qcap = 0;
}

int gga_graph_free(gga_graph *g)
{
    if (!g) return GGA_OK;
    hipSetDevice(g->device);
    if (g->stream) hipStreamSynchronize(g->stream);
    if (g->ms_cache) gga_ms_free(g->ms_cache);
    gga_msbfs_release(g->msbfs);
    dfree(g, (void **)&g->d_row, 4ull * ((uint64_t)g->n + 1));
    dfree(g, (void **)&g->d_col, 4ull * g->m);
    dfree(g, (void **)&g->d_w, 4ull * g->m);
    free_split(g);
    free_queues(g);
    dfree(g, (void **)&g->d_dist, 4ull * g->n);
    dfree(g, (void **)&g->d_mark, 4ull * g->n);
    dfree(g, (void **)&g->d_flag, (uint64_t)g->n);
    dfree(g, (void **)&g->d_ctl, 4ull * GGA_CTL_WORDS);
    dfree(g, (void **)&g->d_cnt, 32ull * GGA_CNT_SLOTS);
    dfree(g, (void **)&g->d_tiles, 4ull * (GGA_TILES((uint64_t)g->n) + 1));
    dfree(g, (void **)&g->d_last, 4ull * g->n);
    dfree(g, (void **)&g->d_tmin, 4ull * GGA_TILES((uint64_t)g->n));
    if (g->h_pin) hipHostFree(g->h_pin);
    for (int i = 0; i < 4; i++)
        if (g->ev[i]) hipEventDestroy(g->ev[i]);
    for (uint32_t i = 0; i < g->kev_cap; i++)
        if (g->kev[i]) hipEventDestroy(g->kev[i]);
    free(g->kev);
    if (g->stream) hipStreamDestroy(g->stream);
    free(g);
    return GGA_OK;
}

static int graph_create(uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *dst,
                        const uint32_t *w, int device, gga_graph **out)
{
    int rc = GGA_OK;
    gga_clear_error();
    if (!out) return GGA_EINVAL;
    *out = NULL;
    if (n == 0 || n == UINT32_MAX) return n ? GGA_ERANGE : GGA_EINVAL;
    if ((uint64_t)n + (uint64_t)m >= UINT32_MAX) {
        gga_set_error("n + m = %llu: chunk ids would exceed uint32", (unsigned long long)n + m);
        return GGA_ERANGE;
    }
    int ndev = gga_device_count();
    if (ndev <= 0) {
        gga_set_error("no HIP device visible");
        return GGA_ENODEV;
    }
    if (device < 0 || device >= ndev) return GGA_EINVAL;
    gga_graph *g = (gga_graph *)calloc(1, sizeof *g);
    if (!g) return GGA_ENOMEM;
    g->device = device;
    g->n = n;
    g->m = m;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (int i = 0; i < 4; i++) HIP_TRY(hipEventCreate(&g->ev[i]));
    HIP_TRY(hipHostMalloc((void **)&g->h_pin, 4ull * GGA_CTL_WORDS, hipHostMallocDefault));
    if ((rc = dmalloc(g, (void **)&g->d_row, 4ull * ((uint64_t)n + 1)))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_col, 4ull * m))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_w, 4ull * m))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_ctl, 4ull * GGA_CTL_WORDS))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_cnt, 32ull * GGA_CNT_SLOTS))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_tiles, 4ull * (GGA_TILES((uint64_t)n) + 1)))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_dist, 4ull * n))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_mark, 4ull * n))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_flag, (uint64_t)n))) goto fail;
    HIP_TRY(hipMemcpyAsync(g->d_row, row_ptr, 4ull * ((uint64_t)n + 1), hipMemcpyHostToDevice, g->stream));
    if (m) HIP_TRY(hipMemcpyAsync(g->d_col, dst, 4ull * m, hipMemcpyHostToDevice, g->stream));
    if (m && w) HIP_TRY(hipMemcpyAsync(g->d_w, w, 4ull * m, hipMemcpyHostToDevice, g->stream));
    else if (m) K_TRY(ggak_fill_u32(g->d_w, 1u, m, g->stream));
    {
        uint64_t ws = 0;
        if (w)
            for (uint32_t e = 0; e < m; e++) ws += w[e];
        else
            ws = m;
        g->wmean = m ? (uint32_t)((ws + m - 1) / m) : 1u;
        if (g->wmean == 0) g->wmean = 1;
    }
    K_TRY(ggak_fill_u32(g->d_ctl, 0u, GGA_CTL_WORDS, g->stream));
    K_TRY(ggak_fill_u32(g->d_mark, 0xFFFFFFFFu, n, g->stream));
    /* max degree + first index on the device (graph.h:338-358) */
    {
        uint64_t *key = (uint64_t *)(g->d_ctl + GGA_CTL_RED);
        K_TRY(ggak_degree_max(g->d_row, n, key, g->stream));
        HIP_TRY(hipMemcpyAsync(g->h_pin, key, 8, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        uint64_t k = (uint64_t)g->h_pin[0] | ((uint64_t)g->h_pin[1] << 32);
        g->maxdeg = (uint32_t)(k >> 32);
        g->maxdeg_idx = k ? 0xFFFFFFFFu - (uint32_t)k : 0u;
        K_TRY(ggak_fill_u32(g->d_ctl, 0u, GGA_CTL_WORDS, g->stream));
    }
    /* validate destinations < n on the host side was the caller's job; check
     * row_ptr shape here cheaply */
    if (row_ptr[0] != 0 || row_ptr[n] != m) {
        gga_set_error("row_ptr[0]=%u row_ptr[n]=%u m=%u", row_ptr[0], row_ptr[n], m);
        rc = GGA_EFORMAT;
        goto fail;
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    *out = g;
    return GGA_OK;
fail:
    gga_graph_free(g);
    return rc;
}

static int check_csr(uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *dst)
{
    if (row_ptr[0] != 0 || row_ptr[n] != m) return GGA_EFORMAT;
    for (uint32_t v = 0; v < n; v++)
        if (row_ptr[v + 1] < row_ptr[v]) {
            gga_set_error("row_ptr not monotone at %u", v);
            return GGA_EFORMAT;
        }
    for (uint32_t e = 0; e < m; e++)
        if (dst[e] >= n) {
            gga_set_error("edge %u: dst %u >= n %u", e, dst[e], n);
            return GGA_EFORMAT;
        }
    return GGA_OK;
}

int gga_graph_from_csr32(uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *dst,
                         const uint32_t *w, int device, gga_graph **out)
{
    if (!out || !row_ptr || (m && !dst)) return GGA_EINVAL;
    if (n == 0) return GGA_EINVAL;
    int rc = check_csr(n, m, row_ptr, dst);
    if (rc) return rc;
    return graph_create(n, m, row_ptr, dst, w, device, out);
}

int gga_graph_from_csr(uint32_t n, uint32_t m, const uint64_t *row_end, const uint32_t *dst,
                       const uint32_t *w, int device, gga_graph **out)
{
    gga_csr c;
    int rc = gga_csr_from_row_end(n, m, row_end, dst, w, &c);
    if (rc) return rc;
    rc = graph_create(c.n, c.m, c.row_ptr, c.dst, w ? c.w : NULL, device, out);
    gga_csr_free(&c);
    return rc;
}

int gga_graph_load_gr(const char *path, int device, gga_graph **out)
{
    gga_csr c;
    int rc = gga_csr_read_gr(path, &c);
    if (rc) return rc;
    rc = graph_create(c.n, c.m, c.row_ptr, c.dst, c.has_weights ? c.w : NULL, device, out);
    gga_csr_free(&c);
    return rc;
}

/* ------------------------------------------------------------------ split */
static int ensure_queues(gga_graph *g, uint32_t cap)
{
    if (g->qcap >= cap && g->d_q0) return GGA_OK;
    free_queues(g);
    int rc = dmalloc(g, (void **)&g->d_q0, 4ull * cap);
    if (!rc) rc = dmalloc(g, (void **)&g->d_q1, 4ull * cap);
    if (rc) {
        free_queues(g);
        return rc;
    }
    g->qcap = cap;
    return GGA_OK;
}

/* gga_graph_split's histogram steps: on failure free the histogram buffers
 * and return without touching the current chunk table */
#define HIST_TRY(expr)                                                                     \
    do {                                                                                   \
        const int e_ = (int)(expr);                                                        \
        if (e_ != 0) {                                                                     \
            gga_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,                       \
                          hipGetErrorString((hipError_t)e_));                              \
            dfree(g, (void **)&hist, 4ull * b);                                            \
            free(h);                                                                       \
            return GGA_EHIP;                                                               \
        }                                                                                  \
    } while (0)

int gga_graph_split(gga_graph *g, uint32_t mda, int32_t level, uint32_t buckets, gga_split_stats *st)
{
    int rc = GGA_OK;
    uint32_t *d_tmp = NULL, *hist = NULL, *h = NULL;
    uint32_t tmp_words = 0, b = 0;
    if (!g) return GGA_EINVAL;
    gga_clear_error();
    HIP_TRY(hipSetDevice(g->device));
    /* multi-source engines read the chunk table from streams of their own:
     * let their work drain before it is rebuilt */
    HIP_TRY(hipDeviceSynchronize());
    uint32_t T;
    uint32_t bucket_rec = 0;
    int32_t level_rec = level;
    if (mda > 0) {
        T = mda;
        level_rec = 0;
    } else if (level > 0) {
        T = (g->maxdeg + (uint32_t)level - 1u) / (uint32_t)level;   /* graph.h:390 */
        if (T == 0) T = 1;                                           /* no edges at all */
    } else if (level < 0) {
        b = buckets ? buckets : 10u;                                 /* bfs_sssp.cu:166 */
        if (b > 4096) return GGA_ERANGE;
        bucket_rec = b;
        rc = dmalloc(g, (void **)&hist, 4ull * b);
        if (rc) return rc;
        h = (uint32_t *)malloc(4ull * b);
        if (!h) {
            dfree(g, (void **)&hist, 4ull * b);
            return GGA_ENOMEM;
        }
        /* a failure in here leaves the previous split in place */
        HIST_TRY(ggak_fill_u32(hist, 0u, b, g->stream));
        HIST_TRY(ggak_degree_hist(g->d_row, g->n, g->maxdeg, b, hist, g->stream));
        HIST_TRY(hipMemcpyAsync(h, hist, 4ull * b, hipMemcpyDeviceToHost, g->stream));
        HIST_TRY(hipStreamSynchronize(g->stream));
        dfree(g, (void **)&hist, 4ull * b);
        /* graph.h:375-384: most populated bucket, first on ties */
        int max = -1, maxindex = -1;
        for (uint32_t i = 0; i < b; i++)
            if ((int)h[i] > max) {
                max = (int)h[i];
                maxindex = (int)i;
            }
        free(h);
        h = NULL;
        maxindex++;
        int maxdeg1 = (int)g->maxdeg + 1;
        T = (uint32_t)((maxdeg1 / (int)b) * maxindex);
        if (T == 0) {
            if (g->maxdeg == 0) T = 1;
            else {
                gga_set_error("histogram threshold is 0 (maxdeg %u, %u buckets)", g->maxdeg, b);
                return GGA_EINVAL;
            }
        }
    } else {
        /* L = 0: no node split (bfs_sssp.cu:170-171); one chunk per vertex */
        T = g->maxdeg ? g->maxdeg : 1u;
    }
    free_split(g);
    g->split_gen++;
    if (g->ms_cache) {
        gga_ms_free(g->ms_cache);
        g->ms_cache = NULL;
    }
    g->have_result = 0;
    tmp_words = ggak_split_tmp_words(g->n);
    if (tmp_words < GGA_MAXCHUNK_SLOTS) tmp_words = GGA_MAXCHUNK_SLOTS;
    if ((rc = dmalloc(g, (void **)&d_tmp, 4ull * tmp_words))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_cstart, 4ull * ((uint64_t)g->n + 1)))) goto fail;
    HIP_TRY(hipEventRecord(g->ev[0], g->stream));
    K_TRY(ggak_split_scan(g->d_row, g->n, T, g->d_cstart, d_tmp, g->stream));
    HIP_TRY(hipMemcpyAsync(g->h_pin, g->d_cstart + g->n, 4, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->C = g->h_pin[0];
    if ((rc = dmalloc(g, (void **)&g->d_cbeg, 4ull * ((uint64_t)g->C + 1)))) goto fail;
    if ((rc = dmalloc(g, (void **)&g->d_cown, 4ull * g->C))) goto fail;
    uint32_t max_chunk = 0;
    {
        uint32_t *d_max = d_tmp;  /* scan scratch is free again */
        K_TRY(ggak_fill_u32(d_max, 0u, GGA_MAXCHUNK_SLOTS, g->stream));
        K_TRY(ggak_split_fill(g->d_row, g->d_cstart, g->n, g->m, g->d_cbeg, g->d_cown, d_max, g->stream));
        HIP_TRY(hipMemcpyAsync(g->h_pin, d_max, 4ull * GGA_MAXCHUNK_SLOTS, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        for (uint32_t i = 0; i < GGA_MAXCHUNK_SLOTS; i++)
            if (g->h_pin[i] > max_chunk) max_chunk = g->h_pin[i];
    }
    g->Cref = g->C;
    /* A chunk longer than GGA_PIECE_MAX edges (a threshold above it that
     * leaves hubs whole: L = 0 / 1, or an explicit large T) would be walked
     * by one wave while the rest of the grid idles -- the load imbalance the
     * reference answers with a child grid per heavy vertex (cudaKernels.h:
     * 466-487).  The kernels then walk a second table whose chunks are the
     * same vertices cut at min(T, GGA_PIECE_MAX) edges, so every heavy
     * vertex is spread over many waves; the reference's table stays for the
     * split statistics and gga_graph_chunks. */
    if (max_chunk > GGA_PIECE_MAX) {
        g->d_rcbeg = g->d_cbeg;
        g->d_rcown = g->d_cown;
        g->d_cbeg = g->d_cown = NULL;
        K_TRY(ggak_split_scan(g->d_row, g->n, GGA_PIECE_MAX, g->d_cstart, d_tmp, g->stream));
        HIP_TRY(hipMemcpyAsync(g->h_pin, g->d_cstart + g->n, 4, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        g->C = g->h_pin[0];
        if ((rc = dmalloc(g, (void **)&g->d_cbeg, 4ull * ((uint64_t)g->C + 1)))) goto fail;
        if ((rc = dmalloc(g, (void **)&g->d_cown, 4ull * g->C))) goto fail;
        K_TRY(ggak_fill_u32(d_tmp, 0u, GGA_MAXCHUNK_SLOTS, g->stream));
        K_TRY(ggak_split_fill(g->d_row, g->d_cstart, g->n, g->m, g->d_cbeg, g->d_cown, d_tmp, g->stream));
    }
    HIP_TRY(hipEventRecord(g->ev[1], g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    dfree(g, (void **)&d_tmp, 4ull * tmp_words);
    g->mda = T;
    g->split = 1;
    memset(&g->sst, 0, sizeof g->sst);
    g->sst.nnode = g->n;
    g->sst.nedge = g->m;
    g->sst.ns_nnode = g->Cref;
    g->sst.ns_nedge = g->m;
    g->sst.extranode = g->Cref - g->n;
    g->sst.maxdeg = g->maxdeg;
    g->sst.level = level_rec;
    g->sst.bucket = bucket_rec;
    g->sst.mda = T;
    g->sst.max_chunk = max_chunk;
    g->sst.exec_chunks = g->C;
    HIP_TRY(hipEventElapsedTime(&g->sst.split_ms, g->ev[0], g->ev[1]));
    if ((rc = ensure_queues(g, g->C))) goto fail;
    if (st) *st = g->sst;
    return GGA_OK;
fail:
    dfree(g, (void **)&d_tmp, 4ull * tmp_words);
    if (rc != GGA_OK) free_split(g);
    return rc;
}

int gga_graph_chunks(gga_graph *g, uint32_t *chunk_begin, uint32_t *owner)
{
    int rc = GGA_OK;
    if (!g) return GGA_EINVAL;
    if (!g->split) return GGA_ENOSPLIT;
    HIP_TRY(hipSetDevice(g->device));
    /* the reference's table (NodeSplit's layout), also where the kernels walk a cut one */
    const uint32_t *cb = g->d_rcbeg ? g->d_rcbeg : g->d_cbeg, *co = g->d_rcown ? g->d_rcown : g->d_cown;
    if (chunk_begin)
        HIP_TRY(hipMemcpyAsync(chunk_begin, cb, 4ull * ((uint64_t)g->Cref + 1), hipMemcpyDeviceToHost, g->stream));
    if (owner) HIP_TRY(hipMemcpyAsync(owner, co, 4ull * g->Cref, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return GGA_OK;
fail:
    return rc;
}

int gga_graph_view(gga_graph *g, gga_graph_view_t *v)
{
    if (!g || !v) return GGA_EINVAL;
    if (!g->split) {
        int rc = gga_graph_split(g, 0, 0, 0, NULL);
        if (rc) return rc;
    }
    v->dg = dev_graph(g);
    v->stream = g->stream;
    v->device = g->device;
    v->split_gen = g->split_gen;
    v->maxdeg = g->maxdeg;
    return GGA_OK;
}

gga_ms **gga_graph_ms_slot(gga_graph *g) { return &g->ms_cache; }

void gga_graph_set_tuned(gga_graph *g, uint32_t grid, uint32_t dense_min, uint32_t layout)
{
    g->tuned_grid = grid;
    g->tuned_dense = dense_min;
    g->tuned_layout = layout;
}

void gga_graph_set_tuned_sssp(gga_graph *g, int32_t mode, uint32_t delta)
{
    g->tuned_mode = mode < 0 ? 0 : mode + 1;
    g->tuned_delta = delta;
}

uint32_t gga_graph_wmean(const gga_graph *g) { return g ? g->wmean : 1u; }

uint32_t gga_graph_tuned_layout(const gga_graph *g) { return g->tuned_layout; }
uint32_t gga_graph_maxdeg(const gga_graph *g) { return g ? g->maxdeg : 0; }
void **gga_graph_msbfs_slot(gga_graph *g) { return &g->msbfs; }

/* -------------------------------------------------------------- traversal */
void gga_opts_init(gga_opts *o)
{
    if (o) memset(o, 0, sizeof *o);
}

static int ensure_kev(gga_graph *g, uint32_t need)
{
    if (need <= g->kev_cap) return GGA_OK;
    uint32_t cap = g->kev_cap ? g->kev_cap : 64;
    while (cap < need) cap *= 2;
    hipEvent_t *nk = (hipEvent_t *)realloc(g->kev, sizeof(hipEvent_t) * cap);
    if (!nk) return GGA_ENOMEM;
    g->kev = nk;
    for (uint32_t i = g->kev_cap; i < cap; i++) {
        if (hipEventCreate(&g->kev[i]) != hipSuccess) {
            g->kev_cap = i;
            return GGA_EHIP;
        }
    }
    g->kev_cap = cap;
    return GGA_OK;
}

static int block_index(uint32_t block) { return block == 128 ? 0 : block == 512 ? 2 : 1; }

static uint32_t wl_grid(gga_graph *g, int persistent, int bfs, int count, uint32_t block)
{
    uint32_t *slot = persistent ? &g->p_grid[bfs][count] : &g->wl_grid[block_index(block)][bfs][count];
    if (*slot == 0) *slot = ggak_wl_grid(g->device, persistent, bfs, count, persistent ? 256u : block);
    return *slot;
}

/* Worklist workgroup size: opts.block, else the tuned one, else 256. */
static uint32_t wl_block(const gga_graph *g, const gga_opts *o)
{
    return o && o->block ? o->block : g->tuned_block ? g->tuned_block : 256u;
}

uint32_t gga_graph_default_grid(gga_graph *g, int bfs, uint32_t block)
{
    return wl_grid(g, 0, bfs ? 1 : 0, 0, block ? block : wl_block(g, NULL));
}

void gga_graph_set_tuned_block(gga_graph *g, uint32_t block) { g->tuned_block = block; }

/* Mesh-like: maximum degree within a few times the mean.  Such graphs have
 * long shortest paths and thin frontiers, so steps are many and small. */
static int mesh_like(const gga_graph *g)
{
    const uint64_t avg = g->n ? ((uint64_t)g->m + g->n - 1) / g->n : 0;
    return g->maxdeg <= 4 * avg + 8;
}

/* AUTO: the tuned mode; else the worklist launches (full occupancy for a
 * few big steps), switching to the persistent kernel once a query on this
 * graph needed more than 256 steps (one launch and a grid barrier per step
 * then beat per-step launches and host checks: measured 2.3x on a pure
 * 1024^2 grid at ~2 400 steps, 0.8x on a shortcut 4096^2 grid at ~160). */
static int choose_mode(const gga_graph *g, int bfs)
{
    if (g->tuned_mode) return g->tuned_mode - 1;
    return g->last_steps[bfs ? 1 : 0] > 256u ? GGA_MODE_PERSISTENT : GGA_MODE_WORKLIST;
}

/* Near-far width when opts leave it 0: the tuned one; else, for SSSP on
 * mesh-like graphs, 32 mean edge weights (pure 1024^2 grid: single source
 * 53 / 51 ms at 8 / 32 mean weights, 32-lane engine 350 / 283 ms); else off. */
static uint32_t auto_delta(const gga_graph *g, int bfs)
{
    if (g->tuned_delta) return g->tuned_delta == 0xFFFFFFFFu ? 0u : g->tuned_delta;
    if (bfs || !mesh_like(g)) return 0u;
    const uint64_t d = 32ull * g->wmean;
    return d < 0xFFFFFFF0ull ? (uint32_t)d : 0xFFFFFFF0u;
}

uint32_t gga_graph_auto_delta(const gga_graph *g, int bfs) { return g ? auto_delta(g, bfs) : 0u; }

/* The engines' default bucket width: half the single-source one (a tuned
 * width is kept).  With waves following their wavefronts, 32 lanes on the
 * 4096^2 grid converged in 312 ms at 16 mean weights against 345 at 32,
 * 473 at 64 (8 lanes: 157 / 165 / 168 ms), while single sources were
 * fastest at 32. */
uint32_t gga_graph_engine_delta(const gga_graph *g)
{
    if (!g) return 0u;
    const uint32_t d = auto_delta(g, 0);
    return g->tuned_delta || d < 2u ? d : d / 2u;
}
int gga_graph_mesh_like(const gga_graph *g) { return g ? mesh_like(g) : 0; }

/* One query on a prepared workspace: dist init, the relax loop of the chosen
 * mode, m_reach, error bits.  `own` = the graph's own workspace: per-launch
 * timing events and the instrumentation buffers are available (they are
 * created or allocated on demand); the static-scratch query runs with
 * neither and allocates nothing.  On return h_pin holds the ctl words and
 * *adv the number of step numbers the query used (the largest ring-slot
 * step number minus the epoch): every stamp it wrote into mark[] is below
 * epoch + adv + 1. */
typedef struct query_plan {
    int bfs, mode, count, own;
    uint32_t src, max_iters, epoch;
} query_plan;

static int run_query(gga_graph *g, gga_dev_graph dg, gga_dev_work dw, const query_plan *p, gga_opts *o,
                     gga_stats *s, uint32_t *adv)
{
    int rc = GGA_OK;
    gga_wg_record *d_tl = NULL;
    uint32_t *d_fin = NULL;
    uint64_t *d_conv = NULL;
    uint64_t tl_bytes = 0, fin_bytes = 0, conv_bytes = 0;
    const int bfs = p->bfs, mode = p->mode, count = p->count, own = p->own;
    /* per-launch timing events (the relax time ktime_ms, the list time
     * ctime_ms) only in counted runs: each event is a barrier packet in the
     * stream, ~5 us of dispatch gap, three per step */
    const int tev = own && count;
    const uint32_t max_iters = p->max_iters;
    float ktime = 0.f;
    int ended = 0;   /* reach pass and control words already fetched */
    *adv = 0;

    HIP_TRY(hipEventRecord(g->ev[0], g->stream));              /* bfs_sssp.cu:350 */
    K_TRY(ggak_query_init(dg, dw, p->src, p->epoch, g->stream));
    if (mode == GGA_MODE_TOPO) {
        /* the reference's loop: changed = false; relax; read changed (bfs_sssp.cu:372-468) */
        if (tev && (rc = ensure_kev(g, 2))) goto fail;
        uint32_t it = 0;
        for (;;) {
            it++;
            g->h_pin[0] = 0;
            HIP_TRY(hipMemcpyAsync(dw.ctl + GGA_CTL_CHANGED, g->h_pin, 4, hipMemcpyHostToDevice, g->stream));
            if (tev) HIP_TRY(hipEventRecord(g->kev[0], g->stream));
            K_TRY(ggak_topo_step(dg, dw, bfs, count, g->stream));
            if (tev) HIP_TRY(hipEventRecord(g->kev[1], g->stream));
            HIP_TRY(hipMemcpyAsync(g->h_pin + 1, dw.ctl + GGA_CTL_CHANGED, 4, hipMemcpyDeviceToHost, g->stream));
            HIP_TRY(hipStreamSynchronize(g->stream));
            if (tev) {
                float t = 0.f;
                HIP_TRY(hipEventElapsedTime(&t, g->kev[0], g->kev[1]));
                ktime += t;
            }
            s->launches++;
            if (!g->h_pin[1]) break;
            if (it >= max_iters) {
                rc = GGA_ETIMEDOUT;
                gga_set_error("topology mode: %u iterations without convergence", it);
                goto fail;
            }
        }
        s->iterations = it;
    } else if (mode == GGA_MODE_WORKLIST) {
        const uint32_t block = wl_block(g, o);
        if (block != 128u && block != 256u && block != 512u) {
            rc = GGA_EINVAL;
            gga_set_error("workgroup size %u: the worklist kernel runs 128, 256 or 512 threads", block);
            goto fail;
        }
        /* a tuned grid belongs to the tuned workgroup size */
        const uint32_t grid = o->grid ? o->grid
                              : (g->tuned_grid && !o->block) ? g->tuned_grid
                                                              : wl_grid(g, 0, bfs, count, block);
        if (grid == 0) {
            rc = GGA_EHIP;
            gga_set_error("occupancy query failed");
            goto fail;
        }
        /* steps per host check: sync_every, else 8 -- but the first batch is
         * one longer than the last query on this graph needed (up to 32; step
         * counts vary by one from run to run), so that repeated queries end
         * after one check and at most two empty launches */
        const uint32_t per = o->sync_every ? o->sync_every : 8u;
        const uint32_t prev = g->last_steps[bfs ? 1 : 0];
        uint32_t nb = o->sync_every ? per : (prev ? (prev < 31u ? prev + 1u : 32u) : per);
        if (tev && (rc = ensure_kev(g, 3 * (nb > per ? nb : per)))) goto fail;
        /* instrumentation buffers (freed below, also on failure) */
        const uint32_t tl_steps = own && o->timeline && o->timeline_cap >= grid ? o->timeline_cap / grid : 0u;
        const uint32_t conv_n = own && o->conv_final && o->conv_rows ? o->conv_cap : 0u;
        tl_bytes = 16ull * tl_steps * grid;
        if (tl_steps && (rc = dmalloc(g, (void **)&d_tl, tl_bytes))) goto fail;
        if (conv_n) {
            fin_bytes = 4ull * g->n;
            conv_bytes = 16ull * conv_n;
            if ((rc = dmalloc(g, (void **)&d_fin, fin_bytes))) goto fail;
            if ((rc = dmalloc(g, (void **)&d_conv, conv_bytes))) goto fail;
            HIP_TRY(hipMemcpyAsync(d_fin, o->conv_final, 4ull * g->n, hipMemcpyHostToDevice, g->stream));
            K_TRY(ggak_fill_u32((uint32_t *)d_conv, 0u, 4ull * conv_n, g->stream));
        }
        uint32_t step = 0;
        for (;;) {
            /* `nb` steps back to back: every kernel reads its list length on
             * the device, so steps after convergence are empty no-ops */
            for (uint32_t k = 0; k < nb; k++) {
                const uint32_t abs_step = step + k;
                gga_wg_record *tl = abs_step < tl_steps ? d_tl + (uint64_t)abs_step * grid : NULL;
                if (tev) HIP_TRY(hipEventRecord(g->kev[3 * k], g->stream));
                K_TRY(ggak_wl_step(dg, dw, abs_step % 6u, bfs, count, grid, block, tl, g->stream));
                if (tev) HIP_TRY(hipEventRecord(g->kev[3 * k + 1], g->stream));
                K_TRY(ggak_compact(dg, dw, abs_step % 6u, g->stream));
                K_TRY(ggak_refill(dg, dw, abs_step % 6u, g->stream));
                if (tev) HIP_TRY(hipEventRecord(g->kev[3 * k + 2], g->stream));
                if (abs_step < conv_n) K_TRY(ggak_conv(dw.dist, d_fin, g->n, d_conv + 2ull * abs_step, g->stream));
            }
            step += nb;
            /* the query's ending (m_reach, counters, all control words) is
             * queued behind every batch, before the host looks: when the
             * batch converged -- repeated queries do at the first look --
             * nothing is left to launch after the host's round trip (~30 us
             * of idle device before the reach pass otherwise) */
            HIP_TRY(hipEventRecord(g->ev[1], g->stream));      /* bfs_sssp.cu:469 */
            K_TRY(ggak_reach(dg, dw, g->stream));
            HIP_TRY(hipMemcpyAsync(g->h_pin, dw.ctl, 4ull * GGA_CTL_WORDS, hipMemcpyDeviceToHost, g->stream));
            HIP_TRY(hipStreamSynchronize(g->stream));
            ended = 1;
            if (tev)
                for (uint32_t k = 0; k < nb; k++) {
                    float t = 0.f, tc = 0.f;
                    HIP_TRY(hipEventElapsedTime(&t, g->kev[3 * k], g->kev[3 * k + 1]));
                    HIP_TRY(hipEventElapsedTime(&tc, g->kev[3 * k + 1], g->kev[3 * k + 2]));
                    ktime += t;
                    s->ctime_ms += tc;
                }
            s->launches += nb;
            nb = per;
            if (g->h_pin[GGA_CTL_TAIL(step % 3u)] == 0 && (!dw.last || g->h_pin[GGA_CTL_FAR] == 0)) break;
            if (step >= max_iters) {
                rc = GGA_ETIMEDOUT;
                gga_set_error("worklist mode: %u steps without convergence", step);
                goto fail;
            }
        }
        if (tl_steps) {
            const uint32_t ns = step < tl_steps ? step : tl_steps;
            HIP_TRY(hipMemcpy(o->timeline, d_tl, 16ull * ns * grid, hipMemcpyDeviceToHost));
            s->timeline_records = ns * grid;
            s->timeline_grid = grid;
        }
        if (conv_n) {
            uint64_t *h = (uint64_t *)malloc(16ull * conv_n);
            if (!h) {
                rc = GGA_ENOMEM;
                goto fail;
            }
            HIP_TRY(hipMemcpy(h, d_conv, 16ull * conv_n, hipMemcpyDeviceToHost));
            const uint32_t ns = step < conv_n ? step : conv_n;
            for (uint32_t k = 0; k < ns; k++) {
                o->conv_rows[k].step = k + 1u;   /* after k + 1 relax steps (bfs_sssp.cu:377, 423) */
                o->conv_rows[k].diff_count = (uint32_t)h[2 * k];
                o->conv_rows[k].norm2 = sqrt((double)h[2 * k + 1]);
            }
            free(h);
            s->conv_steps = ns;
        }
    } else { /* GGA_MODE_PERSISTENT */
        const uint32_t maxg = wl_grid(g, 1, bfs, count, 256u);
        /* half the grid once a query on this graph ran more than 1024 steps:
         * such frontiers are thin, and every step ends in a grid barrier whose
         * cost grows with the grid (pure 4096^2 grid BFS 102 -> 84 ms, SSSP
         * 200 -> 194 ms; 1024^2 25 -> 20 ms, 35 -> 29 ms).  A 4096^2 grid
         * with 0.5% shortcuts (80-125 steps, wide frontiers) needs the full
         * grid: 10 vs 16 ms BFS. */
        const int thin = g->last_steps[bfs ? 1 : 0] > 1024u;
        uint32_t grid = o->grid ? o->grid : (thin && maxg >= 2u ? maxg / 2u : maxg);
        if (grid > maxg) grid = maxg;
        if (grid == 0) {
            rc = GGA_EHIP;
            gga_set_error("occupancy query failed");
            goto fail;
        }
        /* every barrier wait is bounded (5 s); a step never comes near that */
        gga_coop_lock();
        hipError_t he = hipEventRecord(g->ev[2], g->stream);
        int kr = he == hipSuccess ? ggak_persistent(dg, dw, bfs, count, grid, max_iters, 5000u, g->stream) : 0;
        if (he == hipSuccess && kr == 0) he = hipEventRecord(g->ev[3], g->stream);
        const hipError_t hs = hipStreamSynchronize(g->stream);
        gga_coop_unlock();
        HIP_TRY(he);
        K_TRY(kr);
        HIP_TRY(hs);
        HIP_TRY(hipEventElapsedTime(&ktime, g->ev[2], g->ev[3]));
        s->launches = 1;
    }
    if (!ended) {
        HIP_TRY(hipEventRecord(g->ev[1], g->stream));          /* bfs_sssp.cu:469 */
        K_TRY(ggak_reach(dg, dw, g->stream));
        HIP_TRY(hipMemcpyAsync(g->h_pin, dw.ctl, 4ull * GGA_CTL_WORDS, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    HIP_TRY(hipEventElapsedTime(&s->time_ms, g->ev[0], g->ev[1]));
    {
        const uint32_t *c = g->h_pin;
        for (uint32_t i = 0; i < 3; i++) {
            const uint32_t a = c[GGA_CTL_ITER(i)] - p->epoch;
            if (a > *adv) *adv = a;
        }
        if (mode != GGA_MODE_TOPO) {
            s->iterations = c[GGA_CTL_STEPS];
            if (c[GGA_CTL_ERR] & GGA_DEVERR_QFULL) {
                rc = GGA_ERANGE;
                gga_set_error("worklist overflow (capacity %u)", dw.qcap);
                goto fail;
            }
            if (c[GGA_CTL_ERR] & GGA_DEVERR_SPIN) {
                rc = GGA_ETIMEDOUT;
                gga_set_error("persistent kernel: grid barrier spin limit");
                goto fail;
            }
            if (mode == GGA_MODE_PERSISTENT) {
                uint32_t left = c[GGA_CTL_TAIL(s->iterations % 3u)];
                if (left) {
                    rc = GGA_ETIMEDOUT;
                    gga_set_error("persistent kernel: %u steps without convergence", s->iterations);
                    goto fail;
                }
            }
        }
        const uint64_t *c64 = (const uint64_t *)(c + GGA_CTL_CNT);
        s->edges_inspected = c64[0];
        s->chunks_expanded = c64[1];
        s->relaxations = c64[2];
        const uint64_t *r64 = (const uint64_t *)(c + GGA_CTL_RED);
        s->reached = (uint32_t)r64[0];
        s->m_reach = r64[1];
    }
    s->ktime_ms = ktime;
    s->sktime_ms = 0.f;
    g->last_steps[bfs ? 1 : 0] = s->iterations;
    dfree(g, (void **)&d_tl, tl_bytes);
    dfree(g, (void **)&d_fin, fin_bytes);
    dfree(g, (void **)&d_conv, conv_bytes);
    return GGA_OK;
fail:
    hipStreamSynchronize(g->stream);
    dfree(g, (void **)&d_tl, tl_bytes);
    dfree(g, (void **)&d_fin, fin_bytes);
    dfree(g, (void **)&d_conv, conv_bytes);
    return rc;
}

/* Mode, near-far width and step cap of a query, from the options and what
 * the graph has learned (shared by gga_bfs / gga_sssp and the static-scratch
 * query). */
static void plan_query(const gga_graph *g, int bfs, const gga_opts *o, query_plan *p, uint32_t *delta)
{
    p->bfs = bfs;
    p->count = o->count ? 1 : 0;
    int mode = o->mode;
    if (mode == GGA_MODE_AUTO || mode < 0 || mode > GGA_MODE_AUTO) mode = choose_mode(g, bfs);
    p->mode = mode;
    /* Bellman-Ford needs at most n + 1 steps (graph.h:122); near-far adds
     * bucket changes and one empty step per jump over empty buckets, so its
     * default cap is only a guard */
    const uint32_t d = o->delta == 0xFFFFFFFFu ? 0u : (o->delta ? o->delta : auto_delta(g, bfs));
    const int near_far = !bfs && (mode == GGA_MODE_WORKLIST || mode == GGA_MODE_PERSISTENT) && d != 0u;
    *delta = near_far ? d : 0u;
    p->max_iters = o->max_iters ? o->max_iters : (g->n + 1u);
    if (!o->max_iters && near_far)
        p->max_iters = (uint64_t)64u * (g->n + 1u) < 0xFFFFFFF0ull ? 64u * (g->n + 1u) : 0xFFFFFFF0u;
}

static uint32_t dense_default(const gga_graph *g, const gga_opts *o)
{
    return o->dense_min ? o->dense_min
                        : g->tuned_dense ? g->tuned_dense : (g->n / 64u > 4096u ? g->n / 64u : 4096u);
}

static int traverse(gga_graph *g, uint32_t src, int bfs, uint32_t *out, const gga_opts *user,
                    gga_stats *st)
{
    int rc = GGA_OK;
    gga_opts o;
    gga_stats s;
    query_plan p;
    uint32_t delta = 0, adv = 0;
    memset(&s, 0, sizeof s);
    if (user) o = *user;
    else gga_opts_init(&o);
    if (!g) return GGA_EINVAL;
    gga_clear_error();
    if (src >= g->n) {
        gga_set_error("source %u >= n %u", src, g->n);
        return GGA_EINVAL;
    }
    if (!g->split) {
        rc = gga_graph_split(g, 0, 0, 0, NULL);   /* no split: one chunk per vertex */
        if (rc) return rc;
    }
    const uint32_t inf = o.inf ? o.inf : GGA_INF;
    plan_query(g, bfs, &o, &p, &delta);
    p.src = src;
    p.own = 1;
    HIP_TRY(hipSetDevice(g->device));
    if ((rc = ensure_queues(g, g->C))) return rc;
    /* Queue-step stamps in mark[] outlive their query: each query's step
     * numbers start past every stamp written before (g->epoch).  Refill
     * before the numbers wrap, and after a failed query, whose last step
     * number is unknown. */
    if (g->epoch > 0xF0000000u || g->mark_dirty) {
        K_TRY(ggak_fill_u32(g->d_mark, 0xFFFFFFFFu, g->n, g->stream));
        g->epoch = 0;
        g->mark_dirty = 0;
    }
    p.epoch = g->epoch;
    gga_dev_graph dg = dev_graph(g);
    gga_dev_work dw = dev_work(g, inf);
    dw.dense_min = dense_default(g, &o);
    dw.bmin = mesh_like(g) ? GGA_MESH_BMIN : 1u;
    /* near-far buckets: SSSP only (BFS levels are already one bucket per level) */
    if (delta) {
        if (!g->d_last && (rc = dmalloc(g, (void **)&g->d_last, 4ull * g->n))) return rc;
        if (!g->d_tmin && (rc = dmalloc(g, (void **)&g->d_tmin, 4ull * GGA_TILES((uint64_t)g->n)))) return rc;
        dw.last = g->d_last;
        dw.tmin = g->d_tmin;
        dw.delta = delta;
    }
    s.mode = p.mode;
    rc = run_query(g, dg, dw, &p, &o, &s, &adv);
    if (rc) goto fail;
    g->epoch = p.epoch + adv + 1u;
    g->have_result = 1;
    if (out) HIP_TRY(hipMemcpy(out, g->d_dist, 4ull * g->n, hipMemcpyDeviceToHost));   /* bfs_sssp.cu:489 */
    if (st) *st = s;
    return GGA_OK;
fail:
    g->mark_dirty = 1;
    g->have_result = 0;
    if (st) *st = s;
    return rc;
}

int gga_bfs(gga_graph *g, uint32_t src, uint32_t *out, const gga_opts *o, gga_stats *st)
{
    return traverse(g, src, 1, out, o, st);
}

int gga_sssp(gga_graph *g, uint32_t src, uint32_t *out, const gga_opts *o, gga_stats *st)
{
    return traverse(g, src, 0, out, o, st);
}

/* ------------------------------------------------ allocation-free SSSP */
/* Byte offsets of a query's workspace inside the static scratch array
 * (every part 256-byte aligned; the control words keep their 128-byte
 * lines). */
typedef struct scratch_layout {
    uint64_t ctl, cnt, tiles, mark, flag, q0, q1, last, tmin, total;
} scratch_layout;

static uint64_t up256(uint64_t x) { return (x + 255u) & ~255ull; }

static void scratch_plan(uint32_t n, uint32_t C, int near_far, scratch_layout *L)
{
    const uint64_t ntiles = GGA_TILES((uint64_t)n);
    uint64_t o = 0;
    L->ctl = o;
    o = up256(o + 4ull * GGA_CTL_WORDS);
    L->cnt = o;
    o = up256(o + 32ull * GGA_CNT_SLOTS);
    L->tiles = o;
    o = up256(o + 4ull * (ntiles + 1));
    L->mark = o;
    o = up256(o + 4ull * n);
    L->flag = o;
    o = up256(o + (uint64_t)n);
    L->q0 = o;
    o = up256(o + 4ull * C);
    L->q1 = o;
    o = up256(o + 4ull * C);
    L->last = L->tmin = 0;
    if (near_far) {
        L->last = o;
        o = up256(o + 4ull * n);
        L->tmin = o;
        o = up256(o + 4ull * ntiles);
    }
    L->total = o;
}

/* One query at a time uses a device's scratch array. */
static pthread_mutex_t noalloc_mu = PTHREAD_MUTEX_INITIALIZER;

int gga_sssp_noalloc(gga_graph *g, uint32_t src, uint32_t *d_out, const gga_opts *user, gga_stats *st)
{
    int rc = GGA_OK;
    gga_opts o;
    gga_stats s;
    query_plan p;
    scratch_layout L;
    uint32_t delta = 0, adv = 0;
    void *base = NULL;
    memset(&s, 0, sizeof s);
    if (user) o = *user;
    else gga_opts_init(&o);
    if (!g || !d_out) return GGA_EINVAL;
    gga_clear_error();
    if (src >= g->n) {
        gga_set_error("source %u >= n %u", src, g->n);
        return GGA_EINVAL;
    }
    if (!g->split) {
        gga_set_error("gga_sssp_noalloc needs a split graph (gga_graph_split allocates the chunk table)");
        return GGA_ENOSPLIT;
    }
    if (o.timeline || o.conv_rows) {
        gga_set_error("instrumented queries allocate their buffers: not available in gga_sssp_noalloc");
        return GGA_EINVAL;
    }
    const uint32_t inf = o.inf ? o.inf : GGA_INF;
    plan_query(g, 0, &o, &p, &delta);
    p.src = src;
    p.own = 0;
    p.epoch = 0;
    scratch_plan(g->n, g->C, delta != 0, &L);
    if (L.total > GGA_SCRATCH_BYTES) {
        gga_set_error("query workspace of %llu bytes (n = %u, %u chunks%s) exceeds the static scratch of %llu bytes",
                      (unsigned long long)L.total, g->n, g->C, delta ? ", near-far" : "",
                      (unsigned long long)GGA_SCRATCH_BYTES);
        return GGA_ERANGE;
    }
    pthread_mutex_lock(&noalloc_mu);
    HIP_TRY(hipSetDevice(g->device));
    K_TRY(ggak_scratch(g->device, &base));
    {
        uint8_t *b = (uint8_t *)base;
        gga_dev_graph dg = dev_graph(g);
        gga_dev_work dw = dev_work(g, inf);
        dw.dist = d_out;
        dw.mark = (uint32_t *)(b + L.mark);
        dw.flag = b + L.flag;
        dw.q0 = (uint32_t *)(b + L.q0);
        dw.q1 = (uint32_t *)(b + L.q1);
        dw.ctl = (uint32_t *)(b + L.ctl);
        dw.cnt = (uint64_t *)(b + L.cnt);
        dw.tiles = (uint32_t *)(b + L.tiles);
        dw.qcap = g->C;
        dw.dense_min = dense_default(g, &o);
        dw.bmin = mesh_like(g) ? GGA_MESH_BMIN : 1u;
        if (delta) {
            dw.last = (uint32_t *)(b + L.last);
            dw.tmin = (uint32_t *)(b + L.tmin);
            dw.delta = delta;
        }
        /* the array holds whatever the previous call (of any graph) left:
         * the step stamps and every control word start over (fill kernels,
         * no memset nodes) */
        K_TRY(ggak_fill_u32(dw.mark, 0xFFFFFFFFu, g->n, g->stream));
        K_TRY(ggak_fill_u32(dw.ctl, 0u, GGA_CTL_WORDS, g->stream));
        s.mode = p.mode;
        rc = run_query(g, dg, dw, &p, &o, &s, &adv);
    }
    pthread_mutex_unlock(&noalloc_mu);
    if (st) *st = s;
    return rc;
fail:
    hipStreamSynchronize(g->stream);
    pthread_mutex_unlock(&noalloc_mu);
    if (st) *st = s;
    return rc;
}

int gga_graph_copy_dist(gga_graph *g, uint32_t *out)
{
    int rc = GGA_OK;
    if (!g || !out) return GGA_EINVAL;
    if (!g->have_result) return GGA_EINVAL;
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipMemcpy(out, g->d_dist, 4ull * g->n, hipMemcpyDeviceToHost));
    return GGA_OK;
fail:
    return rc;
}

const uint32_t *gga_graph_dist_device(gga_graph *g)
{
    return (g && g->have_result) ? g->d_dist : NULL;
}

int gga_graph_checksum(gga_graph *g, uint64_t *sum)
{
    int rc = GGA_OK;
    if (!g || !sum) return GGA_EINVAL;
    if (!g->have_result) return GGA_EINVAL;
    HIP_TRY(hipSetDevice(g->device));
    gga_dev_graph dg = dev_graph(g);
    gga_dev_work dw = dev_work(g, GGA_INF);
    K_TRY(ggak_checksum(dg, dw, g->stream));
    HIP_TRY(hipMemcpyAsync(g->h_pin, g->d_ctl + GGA_CTL_RED + 4, 8, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    *sum = (uint64_t)g->h_pin[0] | ((uint64_t)g->h_pin[1] << 32);
    return GGA_OK;
fail:
    return rc;
}

int gga_graph_verify(gga_graph *g, uint32_t src, int bfs, uint32_t inf, gga_verify_result *r)
{
    int rc = GGA_OK;
    uint8_t *tight = NULL;
    if (!g || !r) return GGA_EINVAL;
    if (!g->have_result || !g->split) return GGA_EINVAL;
    if (src >= g->n) return GGA_EINVAL;
    if (inf == 0) inf = GGA_INF;
    HIP_TRY(hipSetDevice(g->device));
    if ((rc = dmalloc(g, (void **)&tight, g->n))) return rc;
    HIP_TRY(hipMemsetAsync(tight, 0, g->n, g->stream));
    K_TRY(ggak_fill_u32(g->d_ctl + GGA_CTL_RED, 0u, 8, g->stream));
    gga_dev_graph dg = dev_graph(g);
    gga_dev_work dw = dev_work(g, inf);
    K_TRY(ggak_verify(dg, dw, src, bfs, tight, g->stream));
    HIP_TRY(hipMemcpyAsync(g->h_pin, g->d_ctl + GGA_CTL_RED, 32, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    {
        const uint64_t *r64 = (const uint64_t *)g->h_pin;
        memset(r, 0, sizeof *r);
        r->nerr = r64[0];
        uint64_t bad = r64[1];
        K_TRY(ggak_fill_u32(g->d_ctl + GGA_CTL_RED, 0u, 8, g->stream));
        K_TRY(ggak_reach(dg, dw, g->stream));
        HIP_TRY(hipMemcpyAsync(g->h_pin, g->d_ctl + GGA_CTL_RED, 32, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        uint32_t d0 = 1;
        HIP_TRY(hipMemcpy(&d0, g->d_dist + src, 4, hipMemcpyDeviceToHost));
        r->src_ok = d0 == 0;
        r->untight = bad - (d0 != 0 ? 1u : 0u);
        r->reached = (uint32_t)((const uint64_t *)g->h_pin)[0];
    }
    dfree(g, (void **)&tight, g->n);
    return GGA_OK;
fail:
    dfree(g, (void **)&tight, g->n);
    return rc;
}
