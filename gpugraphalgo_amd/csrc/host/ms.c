/*
 * ms.c -- the 32-lane multi-source engine (include/gga.h, "multi-source
 * engine") and gga_multi_sp.
 *
 * No reference counterpart: the reference runs one source per process
 * (bfs_sssp.cu:203).  Each lane is the single-source relaxation; the engine
 * only shares the steps.  Step bookkeeping: the engine's step number s selects
 * ring slot s % 3 (list lengths, live lanes), list and mask parity s & 1; the
 * kernels take s % 6, which keeps both.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>

#include "gga_host.h"

#ifndef GGA_MS_MESH_BMIN
#define GGA_MS_MESH_BMIN 4u    /* owners per wave batch in engines on mesh-like graphs */
#endif

struct gga_ms {
    gga_graph *g;
    gga_graph_view_t v;
    gga_dev_ms d;
    uint32_t n, ntiles;
    uint32_t step;            /* engine step number (next step to run) */
    uint32_t steps_seen;      /* ctl[STEPS] at the last read-back */
    uint32_t running, done, unit;
    uint32_t grid[2][2];      /* [persistent][count] */
    uint32_t *h_pin;          /* GGA_CTL_WORDS */
    uint32_t *d_stage;        /* 32 x n staging for host fetches (lazy) */
    hipGraphExec_t gx[12];    /* six-step groups captured per phase; [6 + p]: push-only steps */
    hipStream_t own;          /* the engine's stream (v.stream points to it) */
    uint32_t gx_grid[12], gx_dense[12];
    int gx_count[12];
    int mesh;                 /* graph is mesh-like: lists grow a little per step */
    uint32_t last_tail;       /* list length at the last read-back */
    uint32_t dense_default;
    uint64_t dev_bytes;
    hipEvent_t ev[4];         /* [0],[1] around steps; [2] gga_multi_sp's start; [3] read-back done */
    uint32_t pending;         /* steps launched, not yet collected */
    uint32_t launched_lanes;  /* lanes running when the pending steps were launched */
    uint64_t *d_sum, *h_sum;  /* 32 lane fingerprints */
    float ktime_ms;
    uint32_t iterations;
    int failed;               /* a launch reported a device error: lists and barrier
                                 counts are not trusted any more (free the engine) */
};

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

static int ms_alloc(gga_ms *e, void **p, uint64_t bytes)
{
    *p = NULL;
    if (bytes == 0) bytes = 4;
    hipError_t r = hipMalloc(p, bytes);
    if (r != hipSuccess) {
        gga_set_error("hipMalloc %llu bytes: %s", (unsigned long long)bytes, hipGetErrorString(r));
        *p = NULL;
        return GGA_ENOMEM;
    }
    e->dev_bytes += bytes;
    return GGA_OK;
}

static void ms_release(void *p)
{
    if (p) hipFree(p);
}

int gga_ms_free(gga_ms *e)
{
    if (!e) return GGA_OK;
    hipSetDevice(e->v.device);
    if (e->v.stream) hipStreamSynchronize(e->v.stream);
    for (int i = 0; i < 12; i++)
        if (e->gx[i]) hipGraphExecDestroy(e->gx[i]);
    ms_release(e->d.D);
    ms_release(e->d.M0);
    ms_release(e->d.M1);
    ms_release(e->d.q0);
    ms_release(e->d.q1);
    ms_release(e->d.ctl);
    ms_release(e->d.tiles);
    ms_release(e->d.cnt);
    ms_release(e->d_stage);
    if (e->h_pin) hipHostFree(e->h_pin);
    for (int i = 0; i < 4; i++)
        if (e->ev[i]) hipEventDestroy(e->ev[i]);
    ms_release(e->d_sum);
    ms_release(e->d.far);
    ms_release(e->d.lpart);
    ms_release(e->d.tfar);
    if (e->h_sum) hipHostFree(e->h_sum);
    if (e->own) hipStreamDestroy(e->own);
    free(e);
    return GGA_OK;
}

int gga_ms_create(gga_graph *g, uint32_t inf, uint32_t layout, gga_ms **out)
{
    int rc = GGA_OK;
    if (!g || !out) return GGA_EINVAL;
    *out = NULL;
    gga_clear_error();
    gga_ms *e = (gga_ms *)calloc(1, sizeof *e);
    if (!e) return GGA_ENOMEM;
    e->g = g;
    if ((rc = gga_graph_view(g, &e->v))) {
        free(e);
        return rc;
    }
    const gga_dev_graph *dg = &e->v.dg;
    e->n = dg->n;
    e->ntiles = GGA_TILES(dg->n);
    e->d.inf = inf ? inf : GGA_INF;
    e->d.qcap = dg->C;
    if (layout == GGA_MS_AUTO) layout = gga_tune_layout(g);
    e->d.lane_major = layout == GGA_MS_LANE_MAJOR ? 1u : 0u;
    e->dense_default = e->n / 64u > 4096u ? e->n / 64u : 4096u;
    e->d.dense_min = e->dense_default;
    e->mesh = gga_graph_mesh_like(g);
    /* meshes: lanes rarely share a vertex, batches flatten (edge, lane) pairs;
     * at least GGA_MS_MESH_BMIN owners per batch, each wave then following
     * its own wavefront (4096^2 grid, 32 / 8 lanes to convergence: batches
     * of 16, 256 followed pairs: 435 / 173 ms; 4 and 1 024: 343 / 154 ms) */
    e->d.bmin = gga_graph_mesh_like(g) ? GGA_MS_MESH_BMIN : 1u;
    HIP_TRY(hipSetDevice(e->v.device));
    /* a stream of its own, so that engines of one graph run side by side */
    HIP_TRY(hipStreamSynchronize(e->v.stream));
    HIP_TRY(hipStreamCreate(&e->own));
    e->v.stream = e->own;
    for (int i = 0; i < 4; i++) HIP_TRY(hipEventCreate(&e->ev[i]));
    HIP_TRY(hipHostMalloc((void **)&e->h_pin, 4ull * GGA_CTL_WORDS, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&e->h_sum, 8ull * GGA_MS_LANES, hipHostMallocDefault));
    if ((rc = ms_alloc(e, (void **)&e->d_sum, 8ull * GGA_MS_LANES))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.D, 4ull * GGA_MS_LANES * e->n))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.M0, 4ull * e->n))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.M1, 4ull * e->n))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.q0, 4ull * dg->C))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.q1, 4ull * dg->C))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.ctl, 4ull * GGA_CTL_WORDS))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.tiles, 4ull * (e->ntiles + 1ull)))) goto fail;
    if ((rc = ms_alloc(e, (void **)&e->d.cnt, 32ull * GGA_CNT_SLOTS))) goto fail;
    K_TRY(ggak_fill_u32(e->d.D, e->d.inf, (uint64_t)GGA_MS_LANES * e->n, e->v.stream));
    K_TRY(ggak_fill_u32(e->d.M0, 0u, e->n, e->v.stream));
    K_TRY(ggak_fill_u32(e->d.M1, 0u, e->n, e->v.stream));
    K_TRY(ggak_ms_clear(e->d, e->v.stream));
    HIP_TRY(hipStreamSynchronize(e->v.stream));
    if ((rc = gga_ms_set_delta(e, gga_graph_engine_delta(g)))) goto fail;
    *out = e;
    return GGA_OK;
fail:
    gga_ms_free(e);
    return rc;
}

/* Frees the near-far state (far masks, refill partials). */
static void ms_drop_far(gga_ms *e)
{
    if (e->d.far) {
        ms_release(e->d.far);
        e->dev_bytes -= 4ull * e->n;
    }
    if (e->d.lpart) {
        ms_release(e->d.lpart);
        e->dev_bytes -= 4ull * GGA_MS_LANES * GGA_MS_RBLOCKS;
    }
    if (e->d.tfar) {
        ms_release(e->d.tfar);
        e->dev_bytes -= 4ull * GGA_MS_FAR_TILES(e->n);
    }
    e->d.far = NULL;
    e->d.lpart = NULL;
    e->d.tfar = NULL;
}

int gga_ms_set_delta(gga_ms *e, uint32_t delta)
{
    int rc = GGA_OK;
    if (!e) return GGA_EINVAL;
    if (e->running) {
        gga_set_error("near-far width changes only while no lane runs");
        return GGA_EINVAL;
    }
    if (delta == 0xFFFFFFFFu) delta = 0;
    HIP_TRY(hipSetDevice(e->v.device));
    if (delta && !e->d.far) {
        /* the kernels take near-far state from far / lpart being set: both or
         * neither, so any failure below turns near-far off again */
        if ((rc = ms_alloc(e, (void **)&e->d.far, 4ull * e->n))) goto fail;
        if ((rc = ms_alloc(e, (void **)&e->d.lpart, 4ull * GGA_MS_LANES * GGA_MS_RBLOCKS))) goto fail;
        if ((rc = ms_alloc(e, (void **)&e->d.tfar, 4ull * GGA_MS_FAR_TILES(e->n)))) goto fail;
        K_TRY(ggak_fill_u32(e->d.far, 0u, e->n, e->v.stream));
        K_TRY(ggak_fill_u32(e->d.tfar, 0u, GGA_MS_FAR_TILES(e->n), e->v.stream));
        HIP_TRY(hipStreamSynchronize(e->v.stream));
    } else if (!delta && e->d.far) {
        HIP_TRY(hipStreamSynchronize(e->v.stream));
        ms_drop_far(e);
    }
    e->d.delta = delta;
    /* captured steps hold the old engine arguments */
    for (int i = 0; i < 12; i++)
        if (e->gx[i]) {
            hipGraphExecDestroy(e->gx[i]);
            e->gx[i] = NULL;
        }
    return GGA_OK;
fail:
    if (e->d.far && !e->d.delta) {   /* only a half-built near-far state is undone */
        hipStreamSynchronize(e->v.stream);
        ms_drop_far(e);
    }
    return rc;
}

uint64_t gga_ms_graph_bytes(gga_graph *g)
{
    gga_graph_view_t v;
    if (!g || gga_graph_view(g, &v)) return 0;
    uint64_t b = gga_ms_bytes(v.dg.n, v.dg.C);
    if (gga_graph_auto_delta(g, 0)) b += 4ull * v.dg.n + 4ull * GGA_MS_LANES * GGA_MS_RBLOCKS;
    return b;
}

/* Device error bits of the last read-back -> error code; either error
 * leaves the engine's lists (a step interrupted by a barrier timeout) and
 * barrier counts unusable, so the engine refuses further work. */
static int ms_errbits(gga_ms *e)
{
    const uint32_t err = e->h_pin[GGA_CTL_ERR];
    if (!(err & (GGA_DEVERR_QFULL | GGA_DEVERR_SPIN))) return GGA_OK;
    e->failed = 1;
    if (err & GGA_DEVERR_QFULL) {
        gga_set_error("multi-source list overflow (capacity %u)", e->d.qcap);
        return GGA_ERANGE;
    }
    gga_set_error("multi-source persistent kernel: grid barrier spin limit");
    return GGA_ETIMEDOUT;
}

static int ms_check(gga_ms *e)
{
    gga_graph_view_t cur;
    if (e->failed) {
        gga_set_error("engine failed on an earlier device error; free it and create another");
        return GGA_EINVAL;
    }
    int rc = gga_graph_view(e->g, &cur);
    if (rc) return rc;
    if (cur.split_gen != e->v.split_gen) {
        gga_set_error("engine built on an older chunk table (graph re-split)");
        return GGA_EINVAL;
    }
    return GGA_OK;
}

int gga_ms_start(gga_ms *e, uint32_t lanes, const uint32_t *src, uint32_t unit_lanes)
{
    int rc = GGA_OK;
    if (!e || (lanes && !src)) return GGA_EINVAL;
    gga_clear_error();
    if ((rc = ms_check(e))) return rc;
    if (lanes & e->running) {
        gga_set_error("lanes 0x%08x still running", lanes & e->running);
        return GGA_EINVAL;
    }
    gga_ms_sources s;
    memset(&s, 0, sizeof s);
    for (uint32_t i = 0; i < GGA_MS_LANES; i++)
        if ((lanes >> i) & 1u) {
            if (src[i] >= e->n) {
                gga_set_error("lane %u: source %u >= n %u", i, src[i], e->n);
                return GGA_EINVAL;
            }
            s.v[i] = src[i];
        }
    if (!lanes) return GGA_OK;
    HIP_TRY(hipSetDevice(e->v.device));
    e->unit = (e->unit & ~lanes) | (unit_lanes & lanes);
    K_TRY(ggak_ms_start(e->v.dg, e->d, lanes, unit_lanes & lanes, s, e->step % 6u, e->v.stream));
    e->running |= lanes;
    e->done &= ~lanes;
    return GGA_OK;
fail:
    return rc;
}

static uint32_t ms_grid(gga_ms *e, int persistent, int count);
uint32_t gga_ms_step_grid(gga_ms *e) { return e ? ms_grid(e, 0, 0) : 0u; }

static uint32_t ms_grid(gga_ms *e, int persistent, int count)
{
    uint32_t *slot = &e->grid[persistent][count];
    if (*slot == 0) *slot = ggak_ms_grid(e->v.device, persistent, count);
    return *slot;
}

/* Six steps starting at phase p as one captured graph (no memset nodes);
 * push-only steps (no list rebuilds) keep graphs of their own. */
static uint32_t ms_graph_slot(const gga_ms *e, uint32_t p) { return p + (e->d.dense_min == 0xFFFFFFFFu ? 6u : 0u); }

static int ms_graph(gga_ms *e, uint32_t phase, int count, uint32_t grid)
{
    int rc = GGA_OK;
    const uint32_t p = ms_graph_slot(e, phase);
    if (e->gx[p] && e->gx_grid[p] == grid && e->gx_count[p] == count && e->gx_dense[p] == e->d.dense_min)
        return GGA_OK;
    if (e->gx[p]) {
        hipGraphExecDestroy(e->gx[p]);
        e->gx[p] = NULL;
    }
    hipGraph_t gr = NULL;
    HIP_TRY(hipStreamBeginCapture(e->v.stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t k = 0; k < 6; k++) {
        int r = ggak_ms_step(e->v.dg, e->d, (phase + k) % 6u, count, grid, e->v.stream);
        if (r) {
            hipStreamEndCapture(e->v.stream, &gr);
            if (gr) hipGraphDestroy(gr);
            gga_set_error("capture of step %u: %s", k, hipGetErrorString((hipError_t)r));
            return GGA_EHIP;
        }
    }
    HIP_TRY(hipStreamEndCapture(e->v.stream, &gr));
    HIP_TRY(hipGraphInstantiate(&e->gx[p], gr, NULL, NULL, 0));
    hipGraphDestroy(gr);
    e->gx_grid[p] = grid;
    e->gx_count[p] = count;
    e->gx_dense[p] = e->d.dense_min;
    return GGA_OK;
fail:
    if (gr) hipGraphDestroy(gr);
    return rc;
}

/* ctl -> host (synchronous); error bits -> error codes. */
static int ms_readback(gga_ms *e)
{
    int rc = GGA_OK;
    HIP_TRY(hipMemcpyAsync(e->h_pin, e->d.ctl, 4ull * GGA_CTL_WORDS, hipMemcpyDeviceToHost, e->v.stream));
    HIP_TRY(hipStreamSynchronize(e->v.stream));
    return ms_errbits(e);
fail:
    return rc;
}

/* List length and live lanes of the slot the next step reads.  Near-far
 * lanes are live while the device keeps them running (they may wait for a
 * refill with nothing listed). */
static void ms_slot(const gga_ms *e, uint32_t *live, uint32_t *tail)
{
    *tail = e->h_pin[GGA_CTL_TAIL(e->step % 3u)];
    if (e->d.far) *live = e->h_pin[GGA_CTL_RUN];
    else *live = *tail ? e->h_pin[GGA_CTL_LANES(e->step % 3u)] : 0u;
}

int gga_ms_launch(gga_ms *e, uint32_t nsteps, const gga_opts *o)
{
    int rc = GGA_OK;
    gga_opts oo;
    if (!e) return GGA_EINVAL;
    if (e->pending) {
        gga_set_error("previous launch not collected");
        return GGA_EINVAL;
    }
    if ((rc = ms_check(e))) return rc;
    if (o) oo = *o;
    else gga_opts_init(&oo);
    const int count = oo.count ? 1 : 0;
    e->d.dense_min = oo.dense_min ? oo.dense_min : e->dense_default;
    /* Mesh-like graphs grow their lists by a few vertices per lane and step:
     * while the last list was far below the compaction switch, this launch's
     * steps only push, which saves the three list-rebuild launches of every
     * step (a list that does grow past the switch meanwhile is still pushed
     * correctly, only with more appends) */
    if (!oo.dense_min && e->mesh && e->last_tail < e->dense_default / 8u) e->d.dense_min = 0xFFFFFFFFu;
    const uint32_t grid = oo.grid ? oo.grid : ms_grid(e, 0, count);
    if (grid == 0) {
        gga_set_error("occupancy query failed");
        return GGA_EHIP;
    }
    HIP_TRY(hipSetDevice(e->v.device));
    HIP_TRY(hipEventRecord(e->ev[0], e->v.stream));
    uint32_t k = 0;
    if (oo.use_graph)
        for (; k + 6u <= nsteps; k += 6u) {
            if ((rc = ms_graph(e, (e->step + k) % 6u, count, grid))) return rc;
            HIP_TRY(hipGraphLaunch(e->gx[ms_graph_slot(e, (e->step + k) % 6u)], e->v.stream));
        }
    for (; k < nsteps; k++) K_TRY(ggak_ms_step(e->v.dg, e->d, (e->step + k) % 6u, count, grid, e->v.stream));
    HIP_TRY(hipEventRecord(e->ev[1], e->v.stream));
    HIP_TRY(hipMemcpyAsync(e->h_pin, e->d.ctl, 4ull * GGA_CTL_WORDS, hipMemcpyDeviceToHost, e->v.stream));
    HIP_TRY(hipEventRecord(e->ev[3], e->v.stream));
    e->step += nsteps;
    e->pending = nsteps ? nsteps : 1u;   /* a zero-step launch still carries a read-back */
    e->launched_lanes = e->running;
    return GGA_OK;
fail:
    return rc;
}

int gga_ms_collect(gga_ms *e, uint32_t *live_out)
{
    int rc = GGA_OK;
    uint32_t live = 0, tail = 0;
    float t = 0.f;
    if (!e) return GGA_EINVAL;
    if (!e->pending) {
        gga_set_error("nothing launched");
        return GGA_EINVAL;
    }
    HIP_TRY(hipSetDevice(e->v.device));
    HIP_TRY(hipEventSynchronize(e->ev[3]));
    const uint32_t ran = e->pending;
    e->pending = 0;
    if ((rc = ms_errbits(e))) return rc;
    e->steps_seen = e->h_pin[GGA_CTL_STEPS];
    ms_slot(e, &live, &tail);
    e->last_tail = e->h_pin[GGA_CTL_TAIL(e->step % 3u)];
    HIP_TRY(hipEventElapsedTime(&t, e->ev[0], e->ev[1]));
    e->ktime_ms += t;
    e->iterations += ran;
    /* lanes started after the launch are not in this read-back */
    e->done |= e->launched_lanes & ~live;
    e->running &= live | ~e->launched_lanes;
    if (live_out) *live_out = live | (e->running & ~e->launched_lanes);
    return GGA_OK;
fail:
    return rc;
}

int gga_ms_ready(gga_ms *e)
{
    if (!e) return GGA_EINVAL;
    if (!e->pending) return 1;
    hipSetDevice(e->v.device);
    const hipError_t r = hipEventQuery(e->ev[3]);
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) return 0;
    gga_set_error("hipEventQuery: %s", hipGetErrorString(r));
    return GGA_EHIP;
}

int gga_ms_checksum(gga_ms *e, uint32_t lanes, uint64_t *sums)
{
    int rc = GGA_OK;
    if (!e || !sums) return GGA_EINVAL;
    if (lanes & ~e->done) {
        gga_set_error("lanes 0x%08x have not finished", lanes & ~e->done);
        return GGA_EINVAL;
    }
    HIP_TRY(hipSetDevice(e->v.device));
    K_TRY(ggak_fill_u32((uint32_t *)e->d_sum, 0u, 2ull * GGA_MS_LANES, e->v.stream));
    K_TRY(ggak_ms_checksum(e->v.dg, e->d, lanes, e->d_sum, e->v.stream));
    HIP_TRY(hipMemcpyAsync(e->h_sum, e->d_sum, 8ull * GGA_MS_LANES, hipMemcpyDeviceToHost, e->v.stream));
    HIP_TRY(hipStreamSynchronize(e->v.stream));
    for (uint32_t i = 0; i < GGA_MS_LANES; i++) sums[i] = ((lanes >> i) & 1u) ? e->h_sum[i] : 0ull;
    return GGA_OK;
fail:
    return rc;
}

uint64_t gga_ms_bytes(uint32_t n, uint32_t chunks)
{
    return 4ull * GGA_MS_LANES * n + 8ull * n + 8ull * chunks + 4ull * (GGA_TILES((uint64_t)n) + 1) +
           4ull * GGA_CTL_WORDS + 32ull * GGA_CNT_SLOTS + 8ull * GGA_MS_LANES;
}

uint64_t gga_checksum(const uint32_t *dist, uint32_t n)
{
    uint64_t h = 0;
    for (uint32_t v = 0; v < n; v++) h += gga_splitmix64(((uint64_t)v << 32) | dist[v]);
    return h;
}

int gga_ms_run(gga_ms *e, uint32_t max_steps, const gga_opts *o, uint32_t *live_out)
{
    int rc = GGA_OK;
    gga_opts oo;
    if (!e) return GGA_EINVAL;
    gga_clear_error();
    if ((rc = ms_check(e))) return rc;
    if (o) oo = *o;
    else gga_opts_init(&oo);
    const int count = oo.count ? 1 : 0;
    const int persistent = oo.mode == GGA_MODE_PERSISTENT;
    if (persistent && e->d.far) {
        gga_set_error("near-far engines run in step mode (gga_ms_set_delta(e, 0) for the cooperative launch)");
        return GGA_EINVAL;
    }
    /* n + 1 Bellman-Ford steps; near-far adds bucket changes (as gga_sssp) */
    uint32_t cap = max_steps ? max_steps : e->n + 1u;
    if (!max_steps && e->d.far)
        cap = (uint64_t)64u * (e->n + 1u) < 0xFFFFFFF0ull ? 64u * (e->n + 1u) : 0xFFFFFFF0u;
    e->d.dense_min = oo.dense_min ? oo.dense_min : e->dense_default;
    uint32_t live = 0, tail = 0, ran = 0;
    float t = 0.f;
    HIP_TRY(hipSetDevice(e->v.device));
    if (persistent) {
        const uint32_t maxg = ms_grid(e, 1, count);
        uint32_t grid = oo.grid && oo.grid < maxg ? oo.grid : maxg;
        if (grid == 0) {
            gga_set_error("occupancy query failed");
            return GGA_EHIP;
        }
        gga_coop_lock();
        hipError_t he = hipEventRecord(e->ev[0], e->v.stream);
        int kr = he == hipSuccess ? ggak_ms_persistent(e->v.dg, e->d, e->step % 6u, cap, count, grid, 5000u,
                                                       e->v.stream)
                                  : 0;
        if (he == hipSuccess && kr == 0) he = hipEventRecord(e->ev[1], e->v.stream);
        const hipError_t hs = hipStreamSynchronize(e->v.stream);
        gga_coop_unlock();
        HIP_TRY(he);
        K_TRY(kr);
        HIP_TRY(hs);
        rc = ms_readback(e);
        /* the kernel ran exactly the non-empty steps it counted */
        ran = e->h_pin[GGA_CTL_STEPS] - e->steps_seen;
        e->steps_seen = e->h_pin[GGA_CTL_STEPS];
        e->step += ran;
        ms_slot(e, &live, &tail);
        if (rc) return rc;
        HIP_TRY(hipEventElapsedTime(&t, e->ev[0], e->ev[1]));
        e->ktime_ms += t;
    } else {
        const uint32_t per = oo.use_graph ? 6u : (oo.sync_every ? oo.sync_every : 8u);
        for (;;) {
            const uint32_t k = per < cap - ran ? per : cap - ran;
            if ((rc = gga_ms_launch(e, k, &oo))) return rc;
            if ((rc = gga_ms_collect(e, &live))) return rc;
            ran += k;
            tail = e->h_pin[GGA_CTL_TAIL(e->step % 3u)];
            if ((tail == 0 && live == 0) || ran >= cap) break;
        }
        if (live_out) *live_out = live;
        if (!max_steps && (tail || live)) {
            gga_set_error("multi-source: %u steps without convergence", ran);
            return GGA_ETIMEDOUT;
        }
        return GGA_OK;
    }
    e->iterations += ran;
    e->done |= e->running & ~live;
    e->running &= live;
    if (live_out) *live_out = live;
    if (!max_steps && tail) {
        gga_set_error("multi-source: %u steps without convergence", ran);
        return GGA_ETIMEDOUT;
    }
    return GGA_OK;
fail:
    return rc;
}

uint32_t gga_ms_done(const gga_ms *e) { return e ? e->done : 0u; }

int gga_ms_fetch(gga_ms *e, uint32_t lanes, uint32_t *const *out, int on_device)
{
    int rc = GGA_OK;
    if (!e || (lanes && !out)) return GGA_EINVAL;
    gga_clear_error();
    if (lanes & ~e->done) {
        gga_set_error("lanes 0x%08x have not finished", lanes & ~e->done);
        return GGA_EINVAL;
    }
    if (!lanes) return GGA_OK;
    gga_ms_outputs po;
    memset(&po, 0, sizeof po);
    HIP_TRY(hipSetDevice(e->v.device));
    if (!on_device && !e->d_stage)
        if ((rc = ms_alloc(e, (void **)&e->d_stage, 4ull * GGA_MS_LANES * e->n))) return rc;
    for (uint32_t i = 0; i < GGA_MS_LANES; i++)
        if ((lanes >> i) & 1u) {
            if (!out[i]) return GGA_EINVAL;
            po.p[i] = on_device ? out[i] : e->d_stage + (uint64_t)i * e->n;
        }
    K_TRY(ggak_ms_harvest(e->v.dg, e->d, lanes, po, e->v.stream));
    if (!on_device)
        for (uint32_t i = 0; i < GGA_MS_LANES; i++)
            if ((lanes >> i) & 1u)
                HIP_TRY(hipMemcpyAsync(out[i], po.p[i], 4ull * e->n, hipMemcpyDeviceToHost, e->v.stream));
    HIP_TRY(hipStreamSynchronize(e->v.stream));
    return GGA_OK;
fail:
    return rc;
}

int gga_ms_stats(const gga_ms *ce, gga_stats *st)
{
    int rc = GGA_OK;
    gga_ms *e = (gga_ms *)ce;
    if (!e || !st) return GGA_EINVAL;
    memset(st, 0, sizeof *st);
    HIP_TRY(hipSetDevice(e->v.device));
    K_TRY(ggak_ms_fold(e->d, e->v.stream));
    HIP_TRY(hipMemcpyAsync(e->h_pin, e->d.ctl, 4ull * GGA_CTL_WORDS, hipMemcpyDeviceToHost, e->v.stream));
    HIP_TRY(hipStreamSynchronize(e->v.stream));
    {
        const uint64_t *c64 = (const uint64_t *)(e->h_pin + GGA_CTL_CNT);
        st->edges_inspected = c64[0];
        st->chunks_expanded = c64[1];
        st->relaxations = c64[2];
    }
    st->iterations = e->iterations;
    st->ktime_ms = e->ktime_ms;
    st->mode = -1;
    return GGA_OK;
fail:
    return rc;
}

int gga_multi_sp(gga_graph *g, const uint32_t *srcs, uint32_t k, uint32_t unit_lanes, uint32_t *out,
                 const gga_opts *o, gga_stats *st)
{
    int rc = GGA_OK;
    gga_opts oo;
    gga_stats s;
    memset(&s, 0, sizeof s);
    if (!g || !srcs || k == 0 || k > GGA_MS_LANES) return GGA_EINVAL;
    gga_clear_error();
    if (o) oo = *o;
    else gga_opts_init(&oo);
    /* the step kernels run at full occupancy; the cooperative launch is
     * capped at 2 workgroups per CU, too few waves to hide the gathers */
    if (oo.mode != GGA_MODE_PERSISTENT) oo.mode = GGA_MODE_WORKLIST;
    const uint32_t inf = oo.inf ? oo.inf : GGA_INF;
    gga_ms **slot = gga_graph_ms_slot(g);
    const uint32_t want_lm = (oo.ms_layout ? oo.ms_layout : gga_tune_layout(g)) == GGA_MS_LANE_MAJOR;
    if (*slot && ((*slot)->d.inf != inf || (*slot)->running || (*slot)->d.lane_major != want_lm)) {
        gga_ms_free(*slot);
        *slot = NULL;
    }
    if (!*slot && (rc = gga_ms_create(g, inf, want_lm ? GGA_MS_LANE_MAJOR : GGA_MS_VERTEX_MAJOR, slot)))
        return rc;
    gga_ms *e = *slot;
    {
        /* near-far width: opts, else the graph's; the cooperative launch runs
         * plain Bellman-Ford lanes */
        uint32_t want = oo.delta == 0xFFFFFFFFu ? 0u : (oo.delta ? oo.delta : gga_graph_engine_delta(g));
        if (oo.mode == GGA_MODE_PERSISTENT) want = 0;
        if (want != e->d.delta && (rc = gga_ms_set_delta(e, want))) return rc;
    }
    const uint32_t lanes = k == 32 ? 0xFFFFFFFFu : ((1u << k) - 1u);
    uint32_t src[GGA_MS_LANES];
    memset(src, 0, sizeof src);
    memcpy(src, srcs, 4ull * k);
    gga_stats before;
    if (oo.count && (rc = gga_ms_stats(e, &before))) return rc;
    const uint32_t it0 = e->iterations;
    HIP_TRY(hipSetDevice(e->v.device));
    HIP_TRY(hipEventRecord(e->ev[2], e->v.stream));   /* distance init is inside the timed region */
    rc = gga_ms_start(e, lanes, src, unit_lanes);
    if (!rc) rc = gga_ms_run(e, 0, &oo, NULL);
    if (rc) {
        gga_ms_free(e);
        *slot = NULL;
        if (st) *st = s;
        return rc;
    }
    /* ev[1] was recorded after the last step group, before the read-back */
    HIP_TRY(hipEventElapsedTime(&s.time_ms, e->ev[2], e->ev[1]));
    s.iterations = e->iterations - it0;
    s.mode = oo.mode;
    s.launches = oo.mode == GGA_MODE_PERSISTENT ? 1 : s.iterations;
    K_TRY(ggak_ms_reach(e->v.dg, e->d, lanes, e->v.stream));
    {
        gga_stats now;
        if ((rc = gga_ms_stats(e, &now))) return rc;
        const uint64_t *r64 = (const uint64_t *)(e->h_pin + GGA_CTL_RED);
        s.reached = (uint32_t)(r64[0] > 0xFFFFFFFFull ? 0xFFFFFFFFull : r64[0]);
        s.m_reach = r64[1];
        if (oo.count) {
            s.edges_inspected = now.edges_inspected - before.edges_inspected;
            s.chunks_expanded = now.chunks_expanded - before.chunks_expanded;
            s.relaxations = now.relaxations - before.relaxations;
        }
    }
    s.ktime_ms = s.time_ms;
    if (out) {
        uint32_t *ptr[GGA_MS_LANES];
        memset(ptr, 0, sizeof ptr);
        for (uint32_t i = 0; i < k; i++) ptr[i] = out + (uint64_t)i * e->n;
        if ((rc = gga_ms_fetch(e, lanes, ptr, 0))) return rc;
    }
    if (st) *st = s;
    return GGA_OK;
fail:
    if (st) *st = s;
    return rc;
}
