/*
 * serve.c -- the serving loop: BFS/SSSP queries over graphs resident in HBM,
 * admitted by device-memory budget and batched across queries as they arrive.
 *
 * No reference counterpart (the reference answers one query per process,
 * bfs_sssp.cu:103-539).  Each resident graph gets, on demand, up to
 * `engines_per_graph` 32-lane multi-source engines (ms.c); a query occupies
 * one lane from admission to completion.  One round of gga_server_pump:
 *   1. admission, in arrival order: a query takes a free lane of its graph's
 *      fullest engine that has one; with every engine of the graph full (or
 *      none yet) the graph gets another engine if it may hold one more and
 *      the budget allows, after evicting idle engines (least recently used)
 *      if needed; a query whose engine alone exceeds the budget is rejected
 *      (GGA_EBUDGET); a query that cannot be placed waits without blocking
 *      later queries of other graphs;
 *   2. restart of the assigned lanes (per engine, one call);
 *   3. `steps_per_round` steps launched on every engine with running lanes --
 *      each engine owns a stream, so steps of different engines overlap on
 *      the device (thin frontiers of long-diameter graphs leave it mostly
 *      idle otherwise) -- then collected;
 *   4. lanes that converged deliver their result (device fingerprint, and a
 *      copy to the query's buffer if it gave one) and free up for the next
 *      round's admissions.
 * A query's result does not depend on what else runs: lanes are independent
 * (same relaxation, own distances, own convergence).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gga_host.h"

typedef struct srv_engine {
    gga_ms *e;                /* NULL on a dry-run server */
    int resident;
    uint32_t busy;            /* lanes running */
    uint64_t lane_q[GGA_MS_LANES];
    uint32_t dry_left[GGA_MS_LANES];
    double last_used;
    uint32_t start_mask, start_unit;
    uint32_t start_src[GGA_MS_LANES];
    int launched;
} srv_engine;

typedef struct srv_graph {
    gga_graph *g;             /* NULL on a dry-run server */
    uint32_t n, C;
    int mesh;                 /* mesh-like: steps are latency-bound, its engines keep the full grid */
    uint64_t engine_bytes;
    srv_engine eng[GGA_SERVER_MAX_ENGINES];
} srv_graph;

typedef struct srv_query {
    gga_query q;
    gga_query_result r;
} srv_query;

struct gga_server {
    gga_server_opts o;
    srv_graph *gr;
    uint32_t ngr, cap_gr;
    srv_query *qs;
    uint64_t nq, cap_q;
    uint64_t *pend;           /* FIFO of waiting query ids */
    uint64_t npend, cap_pend;
    uint64_t *fin;            /* finished ids not yet polled */
    uint64_t nfin, cap_fin;
    uint64_t running;
    uint64_t used, peak;
    gga_server_counters st;
    struct timespec t0;
    pthread_mutex_t mu;
};

static double mono(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

double gga_server_now(const gga_server *s)
{
    if (!s) return 0.0;
    return mono() - ((double)s->t0.tv_sec + 1e-9 * (double)s->t0.tv_nsec);
}

static int grow(void **p, uint64_t *cap, uint64_t need, size_t elem)
{
    if (need <= *cap) return GGA_OK;
    uint64_t c = *cap ? *cap : 64;
    while (c < need) c *= 2;
    void *np = realloc(*p, c * elem);
    if (!np) return GGA_ENOMEM;
    *p = np;
    *cap = c;
    return GGA_OK;
}

int gga_server_create(const gga_server_opts *o, gga_server **out)
{
    if (!out) return GGA_EINVAL;
    *out = NULL;
    gga_server *s = (gga_server *)calloc(1, sizeof *s);
    if (!s) return GGA_ENOMEM;
    if (o) s->o = *o;
    if (!s->o.hbm_budget) s->o.hbm_budget = 8ull << 30;
    if (!s->o.steps_per_round) s->o.steps_per_round = 12;
    if (!s->o.lanes || s->o.lanes > GGA_MS_LANES) s->o.lanes = GGA_MS_LANES;
    if (!s->o.dry_steps) s->o.dry_steps = 3;
    if (!s->o.engines_per_graph) s->o.engines_per_graph = 1;
    if (s->o.engines_per_graph > GGA_SERVER_MAX_ENGINES) s->o.engines_per_graph = GGA_SERVER_MAX_ENGINES;
    if (s->o.mesh_reserve == 0) s->o.mesh_reserve = 2;
    if (s->o.mesh_reserve < -1 || s->o.mesh_reserve > 6) {
        free(s);
        return GGA_EINVAL;
    }
    clock_gettime(CLOCK_MONOTONIC, &s->t0);
    pthread_mutex_init(&s->mu, NULL);
    *out = s;
    return GGA_OK;
}

int gga_server_free(gga_server *s)
{
    if (!s) return GGA_OK;
    for (uint32_t i = 0; i < s->ngr; i++)
        for (uint32_t k = 0; k < GGA_SERVER_MAX_ENGINES; k++)
            if (s->gr[i].eng[k].e) gga_ms_free(s->gr[i].eng[k].e);
    free(s->gr);
    free(s->qs);
    free(s->pend);
    free(s->fin);
    pthread_mutex_destroy(&s->mu);
    free(s);
    return GGA_OK;
}

static int add_graph(gga_server *s, gga_graph *g, uint32_t n, uint32_t C, uint64_t bytes, uint32_t *id)
{
    uint64_t cap = s->cap_gr;
    int rc = grow((void **)&s->gr, &cap, s->ngr + 1ull, sizeof(srv_graph));
    if (rc) return rc;
    s->cap_gr = (uint32_t)cap;
    srv_graph *x = &s->gr[s->ngr];
    memset(x, 0, sizeof *x);
    x->g = g;
    x->n = n;
    x->C = C;
    x->mesh = g ? gga_graph_mesh_like(g) : 0;
    x->engine_bytes = bytes ? bytes : gga_ms_bytes(n, C);
    if (s->o.host_results) x->engine_bytes += 4ull * GGA_MS_LANES * n;   /* fetch staging */
    if (id) *id = s->ngr;
    s->ngr++;
    return GGA_OK;
}

int gga_server_add_graph(gga_server *s, gga_graph *g, uint32_t *id)
{
    if (!s || !g || s->o.dry_run) return GGA_EINVAL;
    gga_graph_view_t v;
    int rc = gga_graph_view(g, &v);   /* splits with the defaults if needed */
    if (rc) return rc;
    return add_graph(s, g, v.dg.n, v.dg.C, gga_ms_graph_bytes(g), id);
}

int gga_server_add_graph_dry(gga_server *s, uint32_t n, uint32_t chunks, uint32_t *id)
{
    if (!s || !s->o.dry_run || n == 0) return GGA_EINVAL;
    return add_graph(s, NULL, n, chunks, 0, id);
}

int gga_server_submit(gga_server *s, const gga_query *q, uint64_t *id)
{
    int rc;
    if (!s || !q) return GGA_EINVAL;
    if (q->graph >= s->ngr || q->src >= s->gr[q->graph].n) return GGA_EINVAL;
    pthread_mutex_lock(&s->mu);
    if ((rc = grow((void **)&s->qs, &s->cap_q, s->nq + 1, sizeof(srv_query))) ||
        (rc = grow((void **)&s->pend, &s->cap_pend, s->npend + 1, sizeof(uint64_t)))) {
        pthread_mutex_unlock(&s->mu);
        return rc;
    }
    const uint64_t qid = s->nq++;
    srv_query *x = &s->qs[qid];
    memset(x, 0, sizeof *x);
    x->q = *q;
    x->r.id = qid;
    x->r.graph = q->graph;
    x->r.src = q->src;
    x->r.bfs = q->bfs ? 1u : 0u;
    x->r.arrival = q->arrival > 0.0 ? q->arrival : gga_server_now(s);
    x->r.status = 1;   /* waiting */
    s->pend[s->npend++] = qid;
    s->st.submitted++;
    pthread_mutex_unlock(&s->mu);
    if (id) *id = qid;
    return GGA_OK;
}

static void finish_query(gga_server *s, uint64_t qid, int status, double now)
{
    srv_query *x = &s->qs[qid];
    x->r.status = status;
    x->r.done = now;
    if (grow((void **)&s->fin, &s->cap_fin, s->nfin + 1, sizeof(uint64_t)) == GGA_OK) s->fin[s->nfin++] = qid;
    if (status == GGA_OK) s->st.completed++;
    else s->st.rejected++;
}

/* Frees idle engines, least recently used first, until `need` more bytes fit. */
static int make_room(gga_server *s, uint64_t need)
{
    while (s->used + need > s->o.hbm_budget) {
        srv_graph *vg = NULL;
        srv_engine *v = NULL;
        for (uint32_t i = 0; i < s->ngr; i++)
            for (uint32_t k = 0; k < GGA_SERVER_MAX_ENGINES; k++) {
                srv_engine *x = &s->gr[i].eng[k];
                if (!x->resident || x->busy || x->start_mask) continue;
                if (!v || x->last_used < v->last_used) {
                    v = x;
                    vg = &s->gr[i];
                }
            }
        if (!v) return 0;
        if (v->e) gga_ms_free(v->e);
        v->e = NULL;
        v->resident = 0;
        s->used -= vg->engine_bytes;
        s->st.evictions++;
    }
    return 1;
}

/* Layout of a new engine: the options', else the graph's tuned one, else
 * lane-major (also gga_tune_layout's default on mesh-like graphs).  A served engine mostly runs a few lanes, and lane-major rows
 * keep each lane's gathers near one another, where a 128-byte D[v] line
 * carries 32 lanes for the few in use: on the 12-graph serving set, lane-major
 * engines took the p99 at 240 / 330 / 395 q/s offered from 313 / 688 /
 * 736 ms (vertex-major) to 202 / 357 / 492 ms, while 400 queries at once
 * (full 32-lane batches, where vertex-major lines are shared) dropped from
 * 775 to 615 q/s.  GGA_MS_VERTEX_MAJOR in the options restores the latter. */
static uint32_t engine_layout(const gga_server *s, const srv_graph *G)
{
    if (s->o.layout) return s->o.layout;
    const uint32_t t = G->g ? gga_graph_tuned_layout(G->g) : 0u;
    return t ? t : GGA_MS_LANE_MAJOR;
}

/* The engine a query of G goes to: the fullest resident engine with a free
 * lane, else a new one (budget and engines_per_graph permitting), else NULL. */
static srv_engine *place(gga_server *s, srv_graph *G, uint32_t lanes_all, int *rc)
{
    srv_engine *best = NULL;
    uint32_t best_free = 0;
    for (uint32_t k = 0; k < s->o.engines_per_graph; k++) {
        srv_engine *E = &G->eng[k];
        if (!E->resident) continue;
        const uint32_t nfree = (uint32_t)__builtin_popcount(lanes_all & ~(E->busy | E->start_mask));
        if (nfree && (!best || nfree < best_free)) {
            best = E;
            best_free = nfree;
        }
    }
    if (best) return best;
    for (uint32_t k = 0; k < s->o.engines_per_graph; k++) {
        srv_engine *E = &G->eng[k];
        if (E->resident) continue;
        if (!make_room(s, G->engine_bytes)) return NULL;
        if (!s->o.dry_run && (*rc = gga_ms_create(G->g, 0, engine_layout(s, G), &E->e))) return NULL;
        E->resident = 1;
        s->used += G->engine_bytes;
        if (s->used > s->peak) s->peak = s->used;
        s->st.engines_created++;
        return E;
    }
    return NULL;
}

static int admit(gga_server *s, double now)
{
    pthread_mutex_lock(&s->mu);
    uint64_t keep = 0;
    for (uint64_t i = 0; i < s->npend; i++) {
        const uint64_t qid = s->pend[i];
        srv_query *x = &s->qs[qid];
        srv_graph *G = &s->gr[x->q.graph];
        if (G->engine_bytes > s->o.hbm_budget) {
            finish_query(s, qid, GGA_EBUDGET, now);
            continue;
        }
        const uint32_t lanes_all = s->o.lanes == 32 ? 0xFFFFFFFFu : ((1u << s->o.lanes) - 1u);
        int crc = GGA_OK;
        srv_engine *E = place(s, G, lanes_all, &crc);
        if (crc) {
            finish_query(s, qid, crc, now);
            continue;
        }
        if (!E) {
            s->pend[keep++] = qid;
            continue;
        }
        const uint32_t lane = (uint32_t)__builtin_ctz(lanes_all & ~(E->busy | E->start_mask));
        if (E->busy) s->st.joined++;   /* lanes of this engine are running: the new one joins them */
        E->start_mask |= 1u << lane;
        E->start_src[lane] = x->q.src;
        if (x->q.bfs) E->start_unit |= 1u << lane;
        E->lane_q[lane] = qid;
        E->dry_left[lane] = s->o.dry_steps;
        x->r.admitted = now;
        x->r.lane = lane;
        x->r.engine = (uint32_t)(E - G->eng);
        x->r.status = 2;   /* running */
        s->running++;
        if (s->running > s->st.peak_running) s->st.peak_running = (uint32_t)s->running;
    }
    s->npend = keep;
    pthread_mutex_unlock(&s->mu);
    return GGA_OK;
}

/* Results of the converged lanes of engine G. */
static int deliver(gga_server *s, srv_engine *G, uint32_t lanes)
{
    int rc = GGA_OK;
    uint64_t sums[GGA_MS_LANES];
    memset(sums, 0, sizeof sums);
    if (!s->o.dry_run) {
        if ((rc = gga_ms_checksum(G->e, lanes, sums))) return rc;
        uint32_t *hp[GGA_MS_LANES], *dp[GGA_MS_LANES];
        uint32_t hmask = 0, dmask = 0;
        memset(hp, 0, sizeof hp);
        memset(dp, 0, sizeof dp);
        pthread_mutex_lock(&s->mu);   /* a concurrent submit may move qs */
        for (uint32_t l = 0; l < GGA_MS_LANES; l++) {
            if (!((lanes >> l) & 1u)) continue;
            const gga_query *q = &s->qs[G->lane_q[l]].q;
            if (!q->out) continue;
            if (q->out_on_device) {
                dp[l] = q->out;
                dmask |= 1u << l;
            } else {
                hp[l] = q->out;
                hmask |= 1u << l;
            }
        }
        pthread_mutex_unlock(&s->mu);
        if (dmask && (rc = gga_ms_fetch(G->e, dmask, dp, 1))) return rc;
        if (hmask && (rc = gga_ms_fetch(G->e, hmask, hp, 0))) return rc;
    }
    const double now = gga_server_now(s);
    pthread_mutex_lock(&s->mu);
    for (uint32_t l = 0; l < GGA_MS_LANES; l++) {
        if (!((lanes >> l) & 1u)) continue;
        const uint64_t qid = G->lane_q[l];
        s->qs[qid].r.checksum = sums[l];
        finish_query(s, qid, GGA_OK, now);
        s->running--;
    }
    pthread_mutex_unlock(&s->mu);
    G->busy &= ~lanes;
    G->last_used = now;
    return GGA_OK;
}

int gga_server_pump(gga_server *s, uint32_t *active)
{
    int rc = GGA_OK;
    if (!s) return GGA_EINVAL;
    gga_clear_error();
    const double now = gga_server_now(s);
    if ((rc = admit(s, now))) return rc;
    gga_opts o;
    gga_opts_init(&o);
    o.use_graph = s->o.use_graph;
    o.dense_min = s->o.dense_min;
    /* Engines advance independently: one whose steps are done is collected,
     * its converged lanes deliver, its assigned lanes restart and its next
     * steps are launched, while the others keep running. */
    uint32_t act = 0;
    int progressed = 0, inflight = 0;
    for (uint32_t i = 0; i < s->ngr * GGA_SERVER_MAX_ENGINES; i++) {
        srv_engine *G = &s->gr[i / GGA_SERVER_MAX_ENGINES].eng[i % GGA_SERVER_MAX_ENGINES];
        if (!G->resident) continue;
        if (G->launched) {
            if (!s->o.dry_run) {
                const int r = gga_ms_ready(G->e);
                if (r < 0) return r;
                if (r == 0) {
                    inflight = 1;
                    act += (uint32_t)__builtin_popcount(G->busy);
                    continue;
                }
            }
            uint32_t live = 0;
            if (!s->o.dry_run) {
                if ((rc = gga_ms_collect(G->e, &live))) return rc;
            } else {
                for (uint32_t l = 0; l < GGA_MS_LANES; l++)
                    if (((G->busy >> l) & 1u) && G->dry_left[l] > 1) {
                        G->dry_left[l]--;
                        live |= 1u << l;
                    }
            }
            G->launched = 0;
            progressed = 1;
            const uint32_t fin = G->busy & ~live;
            if (fin) {
                if ((rc = deliver(s, G, fin))) return rc;
                if ((rc = admit(s, gga_server_now(s)))) return rc;   /* refill the freed lanes now */
            }
        }
        if (G->start_mask) {
            if (!s->o.dry_run && (rc = gga_ms_start(G->e, G->start_mask, G->start_src, G->start_unit)))
                return rc;
            G->busy |= G->start_mask;
            G->start_mask = 0;
            G->start_unit = 0;
        }
        if (G->busy) {
            const srv_graph *gr = &s->gr[i / GGA_SERVER_MAX_ENGINES];
            o.grid = 0;
            if (!s->o.dry_run && !gr->mesh && s->o.mesh_reserve > 0)
                o.grid = gga_ms_step_grid(G->e) / 8u * (8u - (uint32_t)s->o.mesh_reserve);
            if (!s->o.dry_run && (rc = gga_ms_launch(G->e, s->o.steps_per_round, &o))) return rc;
            G->launched = 1;
            G->last_used = now;
            progressed = 1;
        }
        act += (uint32_t)__builtin_popcount(G->busy);
    }
    if (progressed) s->st.rounds++;
    else if (inflight) {
        struct timespec ts = {0, 20000};   /* nothing ready: let the device run 20 us */
        nanosleep(&ts, NULL);
    }
    if (active) *active = act;
    return GGA_OK;
}

int gga_server_poll(gga_server *s, gga_query_result *res, uint32_t cap, uint32_t *count)
{
    if (!s || !count || (cap && !res)) return GGA_EINVAL;
    pthread_mutex_lock(&s->mu);
    uint32_t k = 0;
    for (; k < cap && k < s->nfin; k++) res[k] = s->qs[s->fin[k]].r;
    memmove(s->fin, s->fin + k, (s->nfin - k) * sizeof(uint64_t));
    s->nfin -= k;
    pthread_mutex_unlock(&s->mu);
    *count = k;
    return GGA_OK;
}

static int idle(gga_server *s)
{
    pthread_mutex_lock(&s->mu);
    const int r = s->npend == 0 && s->running == 0;
    pthread_mutex_unlock(&s->mu);
    return r;
}

int gga_server_drain(gga_server *s)
{
    if (!s) return GGA_EINVAL;
    uint64_t stuck = 0;
    while (!idle(s)) {
        uint32_t act = 0;
        int rc = gga_server_pump(s, &act);
        if (rc) return rc;
        /* waiting queries with nothing running cannot be placed any more */
        stuck = act ? 0 : stuck + 1;
        if (stuck > 2) {
            gga_set_error("queries wait for memory that no running query will free");
            return GGA_EBUDGET;
        }
    }
    return GGA_OK;
}

/* Is qid one of the ascending ids[0..n)? */
static int in_ids(const uint64_t *ids, uint32_t n, uint64_t qid)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (ids[mid] < qid) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && ids[lo] == qid;
}

int gga_server_replay(gga_server *s, const gga_query *trace, uint32_t nq, gga_query_result *res)
{
    int rc = GGA_OK;
    if (!s || (nq && (!trace || !res))) return GGA_EINVAL;
    /* ids of this replay's queries: ascending, and not contiguous when
     * another thread submits meanwhile */
    uint64_t *ids = (uint64_t *)malloc(sizeof(uint64_t) * (nq ? nq : 1));
    if (!ids) return GGA_ENOMEM;
    const double base = gga_server_now(s);
    uint32_t next = 0;
    while (next < nq || !idle(s)) {
        double now = gga_server_now(s) - base;
        for (; next < nq && trace[next].arrival <= now; next++) {
            gga_query q = trace[next];
            q.arrival = base + trace[next].arrival;
            if ((rc = gga_server_submit(s, &q, &ids[next]))) goto out;
        }
        if (idle(s)) {
            if (next < nq) {   /* nothing to do before the next arrival */
                const double wait = trace[next].arrival - (gga_server_now(s) - base);
                if (wait > 0) {
                    struct timespec ts;
                    ts.tv_sec = (time_t)wait;
                    ts.tv_nsec = (long)((wait - (double)ts.tv_sec) * 1e9);
                    nanosleep(&ts, NULL);
                }
            }
            continue;
        }
        uint32_t act = 0;
        if ((rc = gga_server_pump(s, &act))) goto out;
    }
    pthread_mutex_lock(&s->mu);
    for (uint32_t i = 0; i < nq; i++) {
        res[i] = s->qs[ids[i]].r;
        res[i].arrival -= base;
        res[i].admitted -= base;
        res[i].done -= base;
    }
    /* results were delivered here; drop them from the poll queue */
    uint64_t keep = 0;
    for (uint64_t i = 0; i < s->nfin; i++)
        if (!in_ids(ids, nq, s->fin[i])) s->fin[keep++] = s->fin[i];
    s->nfin = keep;
    pthread_mutex_unlock(&s->mu);
out:
    free(ids);
    return rc;
}

int gga_server_replay_plain(gga_server *s, const gga_query *trace, uint32_t nq, int32_t mode,
                            gga_query_result *res)
{
    int rc = GGA_OK;
    if (!s || s->o.dry_run || (nq && (!trace || !res))) return GGA_EINVAL;
    gga_opts o;
    gga_opts_init(&o);
    o.mode = mode;
    const double base = gga_server_now(s);
    for (uint32_t i = 0; i < nq; i++) {
        const gga_query *q = &trace[i];
        if (q->graph >= s->ngr) return GGA_EINVAL;
        double now = gga_server_now(s) - base;
        if (q->arrival > now) {
            const double wait = q->arrival - now;
            struct timespec ts;
            ts.tv_sec = (time_t)wait;
            ts.tv_nsec = (long)((wait - (double)ts.tv_sec) * 1e9);
            nanosleep(&ts, NULL);
        }
        gga_query_result *r = &res[i];
        memset(r, 0, sizeof *r);
        r->id = i;
        r->graph = q->graph;
        r->src = q->src;
        r->bfs = q->bfs ? 1u : 0u;
        r->arrival = q->arrival;
        r->admitted = gga_server_now(s) - base;
        gga_graph *g = s->gr[q->graph].g;
        rc = q->bfs ? gga_bfs(g, q->src, NULL, &o, NULL) : gga_sssp(g, q->src, NULL, &o, NULL);
        if (!rc) rc = gga_graph_checksum(g, &r->checksum);
        if (!rc && q->out) {
            if (q->out_on_device) {
                rc = hipMemcpy(q->out, gga_graph_dist_device(g), 4ull * s->gr[q->graph].n,
                               hipMemcpyDeviceToDevice) == hipSuccess ? GGA_OK : GGA_EHIP;
            } else {
                rc = gga_graph_copy_dist(g, q->out);
            }
        }
        r->status = rc;
        r->done = gga_server_now(s) - base;
        if (rc) return rc;
    }
    return GGA_OK;
}

int gga_server_stats(const gga_server *s, gga_server_counters *st)
{
    if (!s || !st) return GGA_EINVAL;
    *st = s->st;
    st->bytes_in_use = s->used;
    st->peak_bytes = s->peak;
    uint32_t r = 0;
    for (uint32_t i = 0; i < s->ngr; i++)
        for (uint32_t k = 0; k < GGA_SERVER_MAX_ENGINES; k++) r += s->gr[i].eng[k].resident ? 1u : 0u;
    st->engines_resident = r;
    return GGA_OK;
}
