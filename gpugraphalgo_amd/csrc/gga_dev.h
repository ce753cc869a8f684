/*
 * gga_dev.h -- layout shared by the C host layer and the HIP kernels.
 *
 * HBM layout of one resident graph (all uint32, 0-based):
 *   row[n + 1]      CSR offsets                      (reference psrc/noutgoing, graph.h:55)
 *   col[m], w[m]    destinations, weights (SoA, so BFS streams col only)
 *   cstart[n + 1]   first chunk of each vertex; cstart[n] = C
 *                   (= the reference's renumbered parent id v + presum[v], graph.h:453)
 *   cbeg[C + 1]     first edge of each chunk; chunks of a vertex are contiguous
 *                   and consecutive chunks cover consecutive edges, so chunk c
 *                   spans [cbeg[c], cbeg[c + 1])          (reference psrc - 1)
 *   cown[C]         owning original vertex              (reference parentNode)
 * Per-query workspace:
 *   dist[n]         result, indexed by ORIGINAL vertex (no pseudo-children)
 *   mark[n]         last step that queued the vertex (worklist de-duplication)
 *   flag[n]         bytes: low byte of the last flag step that lowered the vertex
 *   q[2][C]         chunk worklists (a vertex enters a list at most once per step)
 *   ctl[]           control words, one 128-byte line per hot counter
 */
#ifndef GGA_DEV_H
#define GGA_DEV_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gga_dev_graph {
    const uint32_t *row;
    const uint32_t *col;
    const uint32_t *w;
    const uint32_t *cstart;
    const uint32_t *cbeg;
    const uint32_t *cown;
    uint32_t n, m, C, mda;
} gga_dev_graph;

typedef struct gga_dev_work {
    uint32_t *dist;
    uint32_t *mark;      /* queue steps: step number of a vertex's last listing (de-duplication) */
    uint8_t *flag;       /* flag steps: low byte of the step that last lowered a vertex (n bytes,
                            set to the byte before the query's first step at query start) */
    uint32_t *q0;
    uint32_t *q1;
    uint32_t *ctl;
    uint64_t *cnt;       /* GGA_CNT_SLOTS x 4 work counters, striped over many addresses */
    uint32_t *tiles;     /* list compaction: per-tile counts, then offsets (GGA_TILES(n) + 1) */
    uint32_t *last;      /* near-far: distance of each vertex at its last expansion (NULL: off) */
    uint32_t *tmin;      /* near-far: per-tile minimum deferred distance (GGA_TILES(n)) */
    uint32_t inf;
    uint32_t qcap;       /* capacity of q0 / q1 in entries (>= C) */
    uint32_t dense_min;  /* a step whose list holds >= dense_min chunks is a flag step */
    uint32_t delta;      /* near-far bucket width (0: plain Bellman-Ford worklist) */
    uint32_t bmin;       /* fewest chunks per wave batch */
} gga_dev_work;

/* Per-workgroup timeline record of an instrumented relax step: start on the
 * 100 MHz device wall clock, duration in its ticks, placement = XCC id << 8 |
 * the SE / SH / CU bits of HW_ID (bits 8..15). */
#ifndef GGA_WG_RECORD_DEFINED
#define GGA_WG_RECORD_DEFINED
typedef struct gga_wg_record {
    uint64_t start;
    uint32_t ticks;
    uint32_t place;
} gga_wg_record;
#endif

/* Multi-source engine: GGA_MS_LANES independent traversals share every step.
 *   D[n * 32]   distances, vertex-major: the 32 lanes of a vertex are one
 *               128-byte line, so relaxing an edge for every lane touches one
 *               line of dist instead of 32 -- the layout for graphs whose
 *               lanes reach the same vertices in the same steps (small-world
 *               graphs); or lane-major (each lane's n distances contiguous),
 *               for graphs whose lanes' frontiers rarely overlap (meshes),
 *               where a vertex line would be fetched for one or two lanes
 *   M[2][n]     lane masks: M[s & 1][v] = lanes whose distance at v dropped in
 *               step s - 1 (step s expands exactly those lanes from v); a
 *               vertex is listed for step s iff its mask is non-zero, and the
 *               finish pass of step s clears the masks it consumed
 *   q0, q1      chunk lists (step s reads q[s & 1])
 *   tiles[GGA_TILES(n) + 1]  compaction counts / offsets */
#define GGA_MS_RBLOCKS 512u   /* workgroups of the near-far refill (rows of lpart) */
#define GGA_MS_FAR_TILES(n) (((n) + 63u) / 64u)
#ifndef GGA_MS_LANES
#define GGA_MS_LANES 32   /* same value as include/gga.h */
#endif
typedef struct gga_dev_ms {
    uint32_t *D;
    uint32_t *M0;
    uint32_t *M1;
    uint32_t *q0;
    uint32_t *q1;
    uint32_t *ctl;       /* GGA_CTL_WORDS words, same layout as gga_dev_work.ctl */
    uint32_t *tiles;
    uint64_t *cnt;       /* GGA_CNT_SLOTS x 4 work counters */
    uint32_t inf;
    uint32_t qcap;
    uint32_t lane_major; /* 1: D[lane * n + v] (each lane contiguous) instead of D[v * 32 + lane] */
    uint32_t dense_min;  /* a step whose list holds >= dense_min chunks rebuilds the next list
                            by compaction of the masks; smaller ones push */
    /* near-far per lane (NULL / 0: plain Bellman-Ford lanes) */
    uint32_t *far;       /* per vertex: lanes whose distance there dropped by a deferred
                            improvement and was not listed or expanded since (n words) */
    uint32_t *tfar;      /* per 64-vertex tile: the lanes that may hold far bits in the tile
                            (ORed with every far bit, rewritten by the refill from the bits
                            it leaves, cleared for restarted lanes; GGA_MS_FAR_TILES(n) words) */
    uint32_t *lpart;     /* refill: per-workgroup per-lane minimum deferred distance */
    uint32_t delta;      /* bucket width */
    uint32_t bmin;       /* fewest chunks per wave batch (more on meshes, whose lanes rarely meet) */
} gga_dev_ms;
typedef struct gga_ms_sources { uint32_t v[GGA_MS_LANES]; } gga_ms_sources;
typedef struct gga_ms_outputs { uint32_t *p[GGA_MS_LANES]; } gga_ms_outputs;

/* Multi-source BFS with bitmask lanes (kernels/msbfs.hip): up to 256 BFS share
 * every level; a vertex holds 4 x 64-bit words per bitmap.
 *   seen / next [n * 4]  visited lanes, lanes reached in the current level
 *   front[(nsnap + 1) * n * 4]  lanes discovered by a level: slot s < nsnap
 *                holds level s (kept to the end), slot nsnap every later level
 *   lev[k * n]   levels, lane-major (row i = source i), written once per
 *                (vertex, lane): at discovery from level nsnap on, else at the
 *                end from the front slots (inf where never reached)
 *   q            push list (forward chunk ids), tiles: its compaction
 *   part[grid * 4]  per-workgroup level partials (frontier vertices,
 *                frontier out-edges, newly reached pairs, their out-edges) */
#define GGA_MSBFS_MAX 256u
#define GGA_MSBFS_SNAP_MAX 16u   /* front slots kept to the end, at most */
typedef struct gga_dev_msbfs {
    uint64_t *seen;
    uint64_t *front;
    uint64_t *next;
    uint32_t *lev;
    uint32_t *q;
    uint32_t *tiles;
    uint32_t *ctl;       /* GGA_CTL_WORDS words */
    uint64_t *part;
    uint32_t k;          /* lanes in use */
    uint32_t inf;
    uint32_t qcap;
    uint32_t alpha;      /* pull (bottom-up) level when frontier out-edges * alpha > m; 0: never */
    uint32_t nsnap;      /* levels whose front slots are kept (0 .. GGA_MSBFS_SNAP_MAX) */
} gga_dev_msbfs;
/* multi-source BFS control words (within the ctl line of CHANGED / STEPS) */
#define GGA_CTL_DIR       (7u * GGA_CTL_LINE + 16u)   /* 0: next level pushes, 1: pulls */
#define GGA_CTL_DONE      (7u * GGA_CTL_LINE + 17u)   /* frontier empty */
#define GGA_CTL_NPULL     (7u * GGA_CTL_LINE + 18u)   /* levels run bottom-up */

/* Device-scope atomics are performed past the per-XCD L2 and same-address
 * ones serialize (about 16 ns each on MI355X), so nothing hot shares one
 * counter: work counters are striped over GGA_CNT_SLOTS addresses. */
#define GGA_CNT_SLOTS 1024u
/* list compaction works on tiles of GGA_TILE vertices */
#define GGA_TILE 2048u
#define GGA_TILES(n) (((n) + GGA_TILE - 1u) / GGA_TILE)

/* ctl word offsets (uint32 index); hot atomics sit on their own 128-byte line */
#define GGA_CTL_LINE      32u
#define GGA_CTL_TAIL(s)   ((s) * GGA_CTL_LINE)          /* s = 0..2: list length of a ring slot */
/* multi-source near-far: line 3 = per-lane thresholds, line 4 = lane sets */
#define GGA_CTL_LTHR(s)   (3u * GGA_CTL_LINE + (s))     /* lane s lists distances < LTHR(s) */
#define GGA_CTL_RUN       (4u * GGA_CTL_LINE)           /* lanes started and not converged */
#define GGA_CTL_RLISTED   (4u * GGA_CTL_LINE + 1u)      /* lanes a refill listed vertices for */
#define GGA_CTL_TDONE     (5u * GGA_CTL_LINE)           /* workgroups done with a step's tail pass */
#define GGA_CTL_ITER(s)   (6u * GGA_CTL_LINE + (s))     /* step number of a ring slot */
#define GGA_CTL_CHANGED   (7u * GGA_CTL_LINE)           /* topology-mode `changed` flag */
#define GGA_CTL_STEPS     (7u * GGA_CTL_LINE + 1u)      /* steps that had a non-empty list */
#define GGA_CTL_ERR       (7u * GGA_CTL_LINE + 2u)      /* error bits (see GGA_DEVERR_*) */
#define GGA_CTL_CNT       (8u * GGA_CTL_LINE)           /* 3 x uint64: edges, chunks, relaxations */
#define GGA_CTL_BAR       (9u * GGA_CTL_LINE)           /* grid barrier: arrivals of group leaders */
#define GGA_CTL_BARGEN    (10u * GGA_CTL_LINE)          /* grid barrier generation */
#define GGA_CTL_RED       (11u * GGA_CTL_LINE)          /* reductions: 4 x uint64 scratch */
#define GGA_BAR_GROUPS    32u                           /* grid barrier: first-level groups */
#define GGA_CTL_BARG(i)   ((12u + (i)) * GGA_CTL_LINE)  /* grid barrier: arrivals within group i */
#define GGA_CTL_WORDS     ((12u + GGA_BAR_GROUPS) * GGA_CTL_LINE)

/* near-far: threshold of a ring slot's list (only vertices with dist < THR
 * are listed); FAR = deferred vertices remain when a refill found none under
 * the next threshold; RTHR = the threshold the current refill lists under */
#define GGA_CTL_THR(s)    (6u * GGA_CTL_LINE + 4u + (s))
#define GGA_CTL_FAR       (6u * GGA_CTL_LINE + 8u)
#define GGA_CTL_RTHR      (6u * GGA_CTL_LINE + 9u)
#define GGA_CTL_RSTEP     (6u * GGA_CTL_LINE + 10u)   /* step number (ITER) of the last refill */
/* non-zero when the count pass after a flag step found a stamped vertex (so
 * the list it rebuilds for ring slot s is not empty) */
#define GGA_CTL_LISTED(s) (6u * GGA_CTL_LINE + 12u + (s))

/* multi-source engine (kernels/ms.hip): lane mask of a ring slot's list */
#define GGA_CTL_LANES(s)  (7u * GGA_CTL_LINE + 8u + (s))
/* multi-source engine: lanes relaxing with weight 1; kept on the device (not a
 * kernel argument) so that steps replayed from a captured graph see restarts */
#define GGA_CTL_UNIT      (7u * GGA_CTL_LINE + 12u)

#define GGA_DEVERR_SPIN   1u   /* barrier spin limit hit */
#define GGA_DEVERR_QFULL  2u   /* worklist overflow (cannot happen with de-duplication) */

#ifdef __cplusplus
}
#endif
#endif
