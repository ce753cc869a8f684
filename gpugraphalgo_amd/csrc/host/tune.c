/*
 * tune.c -- the autotuner: split threshold, launch grid, push/compaction
 * switch and multi-source layout, per graph, by measurement.
 *
 * Replaces the reference's two ways of picking the node-split threshold: the
 * histogram guess getMaxAllowedOutDegree (graph.h:359-385, with the default
 * 10 buckets of bfs_sssp.cu:166) and the hand sweep over NodeSplitLevel in
 * mysc (mysc:20-51).  The histogram threshold is the tuner's first candidate
 * (its seed); every candidate is timed on one query and the fastest kept.
 * The choice never changes a result -- every threshold, workgroup size, grid
 * and switch gives the same distances -- only the time.  The worklist kernel
 * is built for workgroups of 128, 256 and 512 threads (2, 4, 8 waves with a
 * per-wave LDS push buffer each); the tuner times the three at their
 * occupancy grids, then the grid (workgroups resident per CU) for the one
 * it keeps.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>

#include "gga_host.h"

uint32_t gga_tune_layout(gga_graph *g)
{
    if (!g) return GGA_MS_VERTEX_MAJOR;
    const uint32_t t = gga_graph_tuned_layout(g);
    if (t) return t;
    /* Without a measurement: vertex-major on small-world graphs, whose lanes
     * meet at the same vertices in the same steps and share each D[v] line;
     * lane-major on mesh-like graphs, where waves follow single lanes' fronts
     * and each lane's row keeps its gathers close (32 / 8 lanes to
     * convergence: 4096^2 grid 308 / 152 ms lane-major, 352 / 158 ms
     * vertex-major; with 0.5% shortcuts 157 vs 190 ms, 2048^2 12.6 vs 15.2). */
    return gga_graph_mesh_like(g) ? GGA_MS_LANE_MAJOR : GGA_MS_VERTEX_MAJOR;
}

/* 32 sources spread over the vertex range, starting at src. */
static void spread_sources(uint32_t n, uint32_t src, uint32_t *srcs)
{
    for (uint32_t i = 0; i < GGA_MS_LANES; i++) srcs[i] = (uint32_t)((src + (uint64_t)i * (n / GGA_MS_LANES + 1)) % n);
}

/* Best of three runs (device time, ms) of the tuning objective: one query
 * from src, or with GGA_TUNE_MULTI one 32-lane multi-source run. */
static int time_query(gga_graph *g, uint32_t src, int bfs, uint32_t flags, const gga_opts *o, float *ms)
{
    float best = 0.f;
    uint32_t srcs[GGA_MS_LANES];
    if (flags & GGA_TUNE_MULTI) spread_sources(gga_graph_n(g), src, srcs);
    for (int rep = 0; rep < 3; rep++) {
        gga_stats st;
        int rc;
        if (flags & GGA_TUNE_MULTI)
            rc = gga_multi_sp(g, srcs, GGA_MS_LANES, bfs ? 0xFFFFFFFFu : 0u, NULL, o, &st);
        else
            rc = bfs ? gga_bfs(g, src, NULL, o, &st) : gga_sssp(g, src, NULL, o, &st);
        if (rc) return rc;
        if (rep == 0 || st.time_ms < best) best = st.time_ms;
    }
    *ms = best;
    return GGA_OK;
}

static void record(gga_tune_result *r, uint32_t mda, float ms)
{
    if (r->ntried < GGA_TUNE_MAX) {
        r->tried_mda[r->ntried] = mda;
        r->tried_ms[r->ntried] = ms;
    }
    r->ntried++;
}

int gga_tune(gga_graph *g, uint32_t src, uint32_t flags, gga_tune_result *r)
{
    int rc = GGA_OK;
    if (!g || !r) return GGA_EINVAL;
    if (src >= gga_graph_n(g)) return GGA_EINVAL;
    memset(r, 0, sizeof *r);
    r->mode = GGA_MODE_WORKLIST;
    r->delta = 0xFFFFFFFFu;
    const int bfs = (flags & GGA_TUNE_BFS) ? 1 : 0;
    gga_graph_set_tuned(g, 0, 0, 0);
    gga_graph_set_tuned_sssp(g, -1, 0);
    gga_graph_set_tuned_block(g, 0);
    r->block = 256u;
    gga_opts o;
    gga_opts_init(&o);
    o.mode = GGA_MODE_WORKLIST;
    gga_split_stats st;

    /* 1. split threshold: the histogram seed, then a sweep.  On graphs whose
     * maximum degree is below the bucket count the histogram threshold is 0
     * (the reference would divide by it); the unsplit graph seeds instead. */
    const uint32_t maxdeg = gga_graph_maxdeg(g) ? gga_graph_maxdeg(g) : 1u;
    rc = gga_graph_split(g, 0, -1, 10, &st);
    if (rc == GGA_EINVAL) rc = gga_graph_split(g, maxdeg, 0, 0, &st);
    else if (!rc) r->seed_mda = st.mda;
    if (rc) return rc;
    if ((rc = time_query(g, src, bfs, flags, &o, &r->seed_ms))) return rc;
    record(r, st.mda, r->seed_ms);
    const uint32_t seed_t = st.mda;
    uint32_t best_t = seed_t;
    float best = r->seed_ms;
    const uint32_t cand[] = {16u, 32u, 64u, 128u, 256u, 1024u, maxdeg};
    for (size_t i = 0; i < sizeof cand / sizeof cand[0]; i++) {
        const uint32_t t = cand[i] < maxdeg ? cand[i] : maxdeg;
        int seen = t == seed_t;
        for (size_t j = 0; j < i && !seen; j++) seen = (cand[j] < maxdeg ? cand[j] : maxdeg) == t;
        if (seen) continue;
        float ms;
        if ((rc = gga_graph_split(g, t, 0, 0, &st))) return rc;
        if ((rc = time_query(g, src, bfs, flags, &o, &ms))) return rc;
        record(r, t, ms);
        if (ms < best) {
            best = ms;
            best_t = t;
        }
    }
    if ((rc = gga_graph_split(g, best_t, 0, 0, &st))) return rc;
    r->mda = best_t;

    /* 2. grid (workgroups) and 3. the push / compaction switch (single-source
     * objective, worklist launches of plain Bellman-Ford; the multi-source
     * engine sizes its own grid) */
    o.delta = 0xFFFFFFFFu;
    const uint32_t n = gga_graph_n(g);
    if (flags & GGA_TUNE_MULTI) {
        r->best_ms = best;
        r->layout = gga_tune_layout(g);
        goto layout;
    }
    /* workgroup size at its occupancy grid: 128, 256 (timed already: the
     * threshold sweep ran it), 512 threads */
    r->block_ms[1] = best;
    for (uint32_t bi = 0; bi < 3; bi += 2) {
        float ms;
        o.block = bi == 0 ? 128u : 512u;
        if ((rc = time_query(g, src, bfs, flags, &o, &ms))) return rc;
        r->block_ms[bi] = ms;
        if (ms < best) {
            best = ms;
            r->block = o.block;
        }
    }
    o.block = r->block;
    const uint32_t g0 = gga_graph_default_grid(g, bfs, r->block);
    r->grid = g0;
    for (uint32_t div = 2; div <= 4 && g0 / div >= 64u; div *= 2) {
        float ms;
        o.grid = g0 / div;
        if ((rc = time_query(g, src, bfs, flags, &o, &ms))) return rc;
        if (ms < best) {
            best = ms;
            r->grid = o.grid;
        }
    }
    o.grid = r->grid;
    const uint32_t d0 = n / 64u > 4096u ? n / 64u : 4096u;
    r->dense_min = d0;
    const uint32_t dcand[] = {d0 / 4u, d0 * 4u};
    for (size_t i = 0; i < 2; i++) {
        float ms;
        o.dense_min = dcand[i] ? dcand[i] : 1u;
        if ((rc = time_query(g, src, bfs, flags, &o, &ms))) return rc;
        if (ms < best) {
            best = ms;
            r->dense_min = o.dense_min;
        }
    }
    /* 4. launch mode and near-far width: the worklist launches and the
     * persistent kernel, plain Bellman-Ford and buckets of 2, 8, 32 mean
     * weights (SSSP); the two modes (BFS) */
    {
        const uint32_t wm = gga_graph_wmean(g);
        const struct { int32_t mode; uint32_t delta; } cand2[] = {
            {GGA_MODE_WORKLIST, 0xFFFFFFFFu},  {GGA_MODE_PERSISTENT, 0xFFFFFFFFu},
            {GGA_MODE_PERSISTENT, 2u * wm},    {GGA_MODE_PERSISTENT, 8u * wm},
            {GGA_MODE_PERSISTENT, 32u * wm},   {GGA_MODE_WORKLIST, 8u * wm}};
        const size_t nc = bfs ? 2 : sizeof cand2 / sizeof cand2[0];
        r->mode = GGA_MODE_WORKLIST;
        r->delta = 0xFFFFFFFFu;
        float mbest = best;
        for (size_t i = 0; i < nc; i++) {
            gga_opts o2 = o;
            float ms;
            o2.mode = cand2[i].mode;
            o2.delta = cand2[i].delta;
            if (o2.mode == GGA_MODE_PERSISTENT) o2.grid = 0;   /* one workgroup per CU */
            if (i == 0) {
                mbest = best;   /* the configuration timed last above */
                continue;
            }
            if ((rc = time_query(g, src, bfs, flags, &o2, &ms))) return rc;
            if (ms < mbest) {
                mbest = ms;
                r->mode = o2.mode;
                r->delta = o2.delta;
            }
        }
        best = mbest;
        gga_graph_set_tuned_sssp(g, r->mode, bfs ? 0u : r->delta);
    }
    r->best_ms = best;

    r->layout = gga_tune_layout(g);
layout:
    /* 4. multi-source layout: 32 lanes spread over the vertex range */
    if (flags & GGA_TUNE_LAYOUT) {
        uint32_t srcs[GGA_MS_LANES];
        spread_sources(n, src, srcs);
        float lt[2] = {0.f, 0.f};
        for (int l = 0; l < 2; l++) {
            gga_opts mo;
            gga_opts_init(&mo);
            mo.ms_layout = l == 0 ? GGA_MS_VERTEX_MAJOR : GGA_MS_LANE_MAJOR;
            for (int rep = 0; rep < 2; rep++) {
                gga_stats ms;
                if ((rc = gga_multi_sp(g, srcs, GGA_MS_LANES, bfs ? 0xFFFFFFFFu : 0u, NULL, &mo, &ms))) return rc;
                if (rep == 0 || ms.time_ms < lt[l]) lt[l] = ms.time_ms;
            }
        }
        r->layout = lt[1] < lt[0] ? GGA_MS_LANE_MAJOR : GGA_MS_VERTEX_MAJOR;
        r->layout_ms[0] = lt[0];
        r->layout_ms[1] = lt[1];
    }
    gga_graph_set_tuned(g, r->grid, r->dense_min, (flags & GGA_TUNE_LAYOUT) ? r->layout : 0u);
    gga_graph_set_tuned_block(g, r->block);
    return GGA_OK;
}
