/*
 * msbfs.c -- gga_multi_bfs: up to 256 BFS per cooperative launch with bitmask
 * lanes (kernels/msbfs.hip).  No reference counterpart (one BFS per process,
 * bfs_sssp.cu:203); each lane equals gga_bfs from its source.
 *
 * The graph keeps, once built: the transposed CSR (in-edges) with its own
 * chunk table (threshold 32, no chunk for a vertex without in-edges), the
 * per-vertex bitmaps and the level buffer of the largest k seen.  The fronts
 * of the first MSBFS_SNAP levels are kept (32 B per vertex each) so that the
 * level rows are written once at the end; meshes, whose BFS run thousands of
 * levels, keep none.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>

#include "gga_host.h"

#define MSBFS_SNAP 8u

typedef struct msbfs_state {
    int device;
    uint32_t n, m, split_gen;
    /* transposed graph */
    uint32_t *trow, *tcol, *tcstart, *tcbeg, *tcown;
    uint32_t tC, tT;
    float build_ms;
    /* level state */
    uint64_t *seen, *front, *next, *part;
    uint32_t *lev, *q, *tiles, *ctl, *d_src;
    uint64_t lev_words;
    uint32_t qcap, grid, nsnap;
    uint32_t *h_pin;
    hipEvent_t ev[2];
} msbfs_state;

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

static int xalloc(void **p, uint64_t bytes)
{
    *p = NULL;
    hipError_t r = hipMalloc(p, bytes ? bytes : 4);
    if (r != hipSuccess) {
        gga_set_error("hipMalloc %llu bytes: %s", (unsigned long long)bytes, hipGetErrorString(r));
        *p = NULL;
        return GGA_ENOMEM;
    }
    return GGA_OK;
}

static void xfree(void *p)
{
    if (p) hipFree(p);
}

void gga_msbfs_release(void *state)
{
    msbfs_state *s = (msbfs_state *)state;
    if (!s) return;
    hipSetDevice(s->device);
    xfree(s->trow);
    xfree(s->tcol);
    xfree(s->tcstart);
    xfree(s->tcbeg);
    xfree(s->tcown);
    xfree(s->seen);
    xfree(s->front);
    xfree(s->next);
    xfree(s->part);
    xfree(s->lev);
    xfree(s->q);
    xfree(s->tiles);
    xfree(s->ctl);
    xfree(s->d_src);
    if (s->h_pin) hipHostFree(s->h_pin);
    for (int i = 0; i < 2; i++)
        if (s->ev[i]) hipEventDestroy(s->ev[i]);
    free(s);
}

/* Transposed CSR + chunk table with threshold 32 (one-time). */
static int build_transpose(msbfs_state *s, const gga_graph_view_t *v)
{
    int rc = GGA_OK;
    const gga_dev_graph *g = &v->dg;
    uint32_t *indeg = NULL, *tmp = NULL, *pos = NULL, *maxc = NULL;
    const uint32_t n = g->n;
    uint32_t tmp_words = ggak_split_tmp_words(n + 1);
    if (tmp_words < GGA_MAXCHUNK_SLOTS) tmp_words = GGA_MAXCHUNK_SLOTS;
    /* a pull lane walks its in-chunk serially: chunks of at most 32 in-edges
     * keep the lanes of a wave balanced, whatever the forward threshold (256
     * BFS on R-MAT s20: 8 / 16 / 32 / 64 / 128 give 2.06 / 1.90 / 1.88 /
     * 1.95 / 2.14 ms) */
    s->tT = 32u;
    if ((rc = xalloc((void **)&indeg, 4ull * n))) goto fail;
    if ((rc = xalloc((void **)&pos, 4ull * (n + 1ull)))) goto fail;
    if ((rc = xalloc((void **)&tmp, 4ull * tmp_words))) goto fail;
    if ((rc = xalloc((void **)&s->trow, 4ull * (n + 1ull)))) goto fail;
    if ((rc = xalloc((void **)&s->tcol, 4ull * g->m))) goto fail;
    if ((rc = xalloc((void **)&s->tcstart, 4ull * (n + 1ull)))) goto fail;
    HIP_TRY(hipEventRecord(s->ev[0], v->stream));
    K_TRY(ggak_fill_u32(indeg, 0u, n, v->stream));
    K_TRY(ggak_indeg(g->col, g->m, indeg, v->stream));
    K_TRY(ggak_scan_u32(indeg, n, s->trow, tmp, v->stream));
    HIP_TRY(hipMemcpyAsync(pos, s->trow, 4ull * (n + 1ull), hipMemcpyDeviceToDevice, v->stream));
    K_TRY(ggak_transpose_fill(*g, pos, s->tcol, v->stream));
    /* vertices without in-edges get no in-chunk: a pull level has nothing to
     * look at there (R-MAT s20: half the vertices) */
    K_TRY(ggak_split_scan_ex(s->trow, n, s->tT, 0u, s->tcstart, tmp, v->stream));
    HIP_TRY(hipMemcpyAsync(s->h_pin, s->tcstart + n, 4, hipMemcpyDeviceToHost, v->stream));
    HIP_TRY(hipStreamSynchronize(v->stream));
    s->tC = s->h_pin[0];
    if ((rc = xalloc((void **)&s->tcbeg, 4ull * (s->tC + 1ull)))) goto fail;
    if ((rc = xalloc((void **)&s->tcown, 4ull * s->tC))) goto fail;
    maxc = tmp;
    K_TRY(ggak_fill_u32(maxc, 0u, GGA_MAXCHUNK_SLOTS, v->stream));
    K_TRY(ggak_split_fill(s->trow, s->tcstart, n, g->m, s->tcbeg, s->tcown, maxc, v->stream));
    HIP_TRY(hipEventRecord(s->ev[1], v->stream));
    HIP_TRY(hipStreamSynchronize(v->stream));
    HIP_TRY(hipEventElapsedTime(&s->build_ms, s->ev[0], s->ev[1]));
fail:
    xfree(indeg);
    xfree(pos);
    xfree(tmp);
    return rc;
}

static int get_state(gga_graph *gr, const gga_graph_view_t *v, uint32_t k, msbfs_state **out)
{
    int rc = GGA_OK;
    void **slot = gga_graph_msbfs_slot(gr);
    msbfs_state *s = (msbfs_state *)*slot;
    if (s && s->split_gen != v->split_gen) {   /* chunk tables changed */
        gga_msbfs_release(s);
        *slot = s = NULL;
    }
    if (!s) {
        s = (msbfs_state *)calloc(1, sizeof *s);
        if (!s) return GGA_ENOMEM;
        s->device = v->device;
        s->n = v->dg.n;
        s->m = v->dg.m;
        s->split_gen = v->split_gen;
        *slot = s;
        for (int i = 0; i < 2; i++) HIP_TRY(hipEventCreate(&s->ev[i]));
        HIP_TRY(hipHostMalloc((void **)&s->h_pin, 4ull * GGA_CTL_WORDS, hipHostMallocDefault));
        if ((rc = build_transpose(s, v))) goto fail;
        const uint64_t words = 4ull * s->n;
        if ((rc = xalloc((void **)&s->seen, 8ull * words))) goto fail;
        s->nsnap = gga_graph_mesh_like(gr) ? 0u : MSBFS_SNAP;
        if ((rc = xalloc((void **)&s->front, 8ull * words * (s->nsnap + 1u)))) goto fail;
        if ((rc = xalloc((void **)&s->next, 8ull * words))) goto fail;
        s->qcap = v->dg.C;
        if ((rc = xalloc((void **)&s->q, 4ull * s->qcap))) goto fail;
        if ((rc = xalloc((void **)&s->tiles, 4ull * (GGA_TILES((uint64_t)s->n) + 1)))) goto fail;
        if ((rc = xalloc((void **)&s->ctl, 4ull * GGA_CTL_WORDS))) goto fail;
        if ((rc = xalloc((void **)&s->d_src, 4ull * GGA_MSBFS_MAX))) goto fail;
        K_TRY(ggak_fill_u32(s->ctl, 0u, GGA_CTL_WORDS, v->stream));
        /* more workgroups per CU hide more gather latency (256 BFS on R-MAT
         * s20: 4.8 ms with one, 3.9 with two, before the kept fronts; then
         * 2.43 with two, 2.36 with three, 2.44 with four); meshes run
         * thousands of levels, where the grid barriers of a larger grid cost
         * more (512^2: 148 ms with one, 156 with two) */
        s->grid = ggak_msbfs_grid(v->device, gga_graph_mesh_like(gr) ? 1u : 3u);
        if (s->grid == 0) {
            gga_set_error("occupancy query failed");
            rc = GGA_EHIP;
            goto fail;
        }
        if ((rc = xalloc((void **)&s->part, 32ull * s->grid))) goto fail;
    }
    const uint64_t need = (uint64_t)k * s->n;
    if (need > s->lev_words) {
        xfree(s->lev);
        s->lev = NULL;
        s->lev_words = 0;
        if ((rc = xalloc((void **)&s->lev, 4ull * need))) goto fail;
        s->lev_words = need;
    }
    *out = s;
    return GGA_OK;
fail:
    gga_msbfs_release(s);
    *slot = NULL;
    return rc;
}

int gga_multi_bfs(gga_graph *g, const uint32_t *srcs, uint32_t k, uint32_t *out, const gga_opts *o,
                  gga_stats *st)
{
    int rc = GGA_OK;
    gga_opts oo;
    gga_stats r;
    gga_graph_view_t v;
    memset(&r, 0, sizeof r);
    if (!g || !srcs || k == 0 || k > GGA_MSBFS_MAX) return GGA_EINVAL;
    gga_clear_error();
    if (o) oo = *o;
    else gga_opts_init(&oo);
    if ((rc = gga_graph_view(g, &v))) return rc;
    for (uint32_t i = 0; i < k; i++)
        if (srcs[i] >= v.dg.n) {
            gga_set_error("source %u = %u >= n %u", i, srcs[i], v.dg.n);
            return GGA_EINVAL;
        }
    HIP_TRY(hipSetDevice(v.device));
    msbfs_state *s = NULL;
    if ((rc = get_state(g, &v, k, &s))) return rc;
    gga_dev_msbfs b;
    memset(&b, 0, sizeof b);
    b.seen = s->seen;
    b.front = s->front;
    b.next = s->next;
    b.lev = s->lev;
    b.q = s->q;
    b.tiles = s->tiles;
    b.ctl = s->ctl;
    b.part = s->part;
    b.k = k;
    b.inf = oo.inf ? oo.inf : GGA_INF;
    b.qcap = s->qcap;
    b.alpha = oo.pull_alpha == 0 ? 14u : (oo.pull_alpha == 0xFFFFFFFFu ? 0u : oo.pull_alpha);
    b.nsnap = s->nsnap;
    gga_dev_graph t;
    memset(&t, 0, sizeof t);
    t.row = s->trow;
    t.col = s->tcol;
    t.w = NULL;
    t.cstart = s->tcstart;
    t.cbeg = s->tcbeg;
    t.cown = s->tcown;
    t.n = v.dg.n;
    t.m = v.dg.m;
    t.C = s->tC;
    t.mda = s->tT;
    HIP_TRY(hipMemcpyAsync(s->d_src, srcs, 4ull * k, hipMemcpyHostToDevice, v.stream));
    HIP_TRY(hipEventRecord(s->ev[0], v.stream));
    K_TRY(ggak_msbfs_init(v.dg, b, s->d_src, v.stream));
    {
        /* every barrier wait is bounded (5 s) */
        gga_coop_lock();
        int kr = ggak_msbfs_run(v.dg, t, b, v.dg.n + 1u, s->grid, 5000u, v.stream);
        hipError_t he = kr == 0 ? hipEventRecord(s->ev[1], v.stream) : hipSuccess;
        if (kr == 0 && he == hipSuccess)
            he = hipMemcpyAsync(s->h_pin, s->ctl, 4ull * GGA_CTL_WORDS, hipMemcpyDeviceToHost, v.stream);
        const hipError_t hs = hipStreamSynchronize(v.stream);
        gga_coop_unlock();
        K_TRY(kr);
        HIP_TRY(he);
        HIP_TRY(hs);
    }
    HIP_TRY(hipEventElapsedTime(&r.time_ms, s->ev[0], s->ev[1]));
    {
        const uint32_t *c = s->h_pin;
        /* a launch that left a barrier early or overflowed leaves state the
         * next launch must not inherit: drop the cached state (rebuilt on
         * the next call) */
        if (c[GGA_CTL_ERR] & GGA_DEVERR_SPIN) {
            gga_set_error("multi-source BFS: grid barrier spin limit");
            rc = GGA_ETIMEDOUT;
            gga_msbfs_release(s);
            *gga_graph_msbfs_slot(g) = NULL;
            goto fail;
        }
        if (c[GGA_CTL_ERR] & GGA_DEVERR_QFULL) {
            gga_set_error("multi-source BFS: push list overflow");
            rc = GGA_ERANGE;
            gga_msbfs_release(s);
            *gga_graph_msbfs_slot(g) = NULL;
            goto fail;
        }
        if (!c[GGA_CTL_DONE]) {
            gga_set_error("multi-source BFS: level cap reached");
            rc = GGA_ETIMEDOUT;
            goto fail;
        }
        const uint64_t *tot = (const uint64_t *)(c + GGA_CTL_CNT);
        r.iterations = c[GGA_CTL_STEPS];
        r.pull_levels = c[GGA_CTL_NPULL];
        r.reached = (uint32_t)(tot[0] > 0xFFFFFFFFull ? 0xFFFFFFFFull : tot[0]);
        r.m_reach = tot[1];
    }
    r.ktime_ms = r.time_ms;
    r.mode = GGA_MODE_PERSISTENT;
    r.launches = 1;
    if (out) HIP_TRY(hipMemcpy(out, s->lev, 4ull * k * v.dg.n, hipMemcpyDeviceToHost));
    if (st) *st = r;
    return GGA_OK;
fail:
    if (st) *st = r;
    return rc;
}
