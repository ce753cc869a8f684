/*
 * gga.h -- C-ABI of the MI355X-native BFS/SSSP-with-node-splitting library.
 *
 * The reference (vivekkumark/GPUGraphAlgo, Programs/CUDA_src/Kepler_BFS_SSSP)
 * has no library API: its boundary is the executable
 *     na_{bfs,sssp}.exe <g.gr> [L] [bucket | outfile]     (bfs_sssp.cu:149-185)
 * built from graph.h (CSR model, .gr reader, NodeSplit, serial CPU BFS/SSSP),
 * cudaKernels.h (relax kernels) and myutils.h (verify + output writers).
 * Every entry point below names the reference function or code region it
 * replaces.  All arithmetic is unsigned 32-bit integer.  Plain pointers and
 * sizes only; no C++ or torch types cross this boundary.
 *
 * Conventions
 *   - Vertex ids and edge offsets are uint32 (the reference's `unsigned`,
 *     graph.h:54-55).  Edge arrays are 0-based here (the reference keeps a
 *     null slot 0 and 1-based psrc, graph.h:251-254).
 *   - Distances: uint32; unreached vertices hold `inf` (GGA_INF by default,
 *     GGA_REF_INF = the reference's MYINFINITY 1e9, bfs_sssp.cu:53).
 *   - BFS is SSSP with every weight equal to 1 (cudaKernels.h:320, graph.h:163).
 *   - Every function returns GGA_OK (0) or a negative GGA_E* code; nothing in
 *     the library calls exit() (the reference's CudaTest does, graph.h:523-533).
 *   - A gga_graph owns one HIP stream on its device.  Calls on one handle must
 *     be serialised by the caller; distinct handles may run concurrently.
 */
#ifndef GGA_H
#define GGA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
#define GGA_OK          0
#define GGA_EINVAL     (-1)  /* bad argument: NULL pointer, source >= n, level == 0 ... */
#define GGA_EIO        (-2)  /* open / stat / mmap / read / write failure */
#define GGA_EFORMAT    (-3)  /* malformed .gr: version != 1, truncated, dst >= n, edge size not 0 or 4 */
#define GGA_ERANGE     (-4)  /* counts beyond uint32 ids / offsets, or beyond a kernel limit */
#define GGA_ENOMEM     (-5)  /* host or device allocation failed */
#define GGA_EHIP       (-6)  /* HIP runtime error (message in gga_last_error()) */
#define GGA_ENOSPLIT   (-7)  /* traversal requested before gga_graph_split */
#define GGA_ETIMEDOUT  (-8)  /* iteration cap or persistent-kernel barrier spin limit reached */
#define GGA_EBUDGET    (-9)  /* serving: request does not fit the device-memory budget */
#define GGA_ENODEV    (-10)  /* no HIP device */

#define GGA_INF       0xFFFFFFFFu
#define GGA_REF_INF   1000000000u   /* MYINFINITY, bfs_sssp.cu:53 / graph.h:96 */

/* Returns a static string for an error code. */
const char *gga_strerror(int rc);
/* Last HIP / IO error detail recorded by this thread ("" if none). */
const char *gga_last_error(void);
/* Library version string. */
const char *gga_version(void);
/* Number of visible HIP devices (0 on a host without GPUs; never fails). */
int gga_device_count(void);

/* ------------------------------------------------------- host CSR graphs */
typedef struct gga_csr {
    uint32_t n;           /* vertices */
    uint32_t m;           /* directed edges */
    uint32_t *row_ptr;    /* n + 1 offsets, row_ptr[0] = 0, row_ptr[n] = m */
    uint32_t *dst;        /* m destinations */
    uint32_t *w;          /* m weights (all 1 when the source carried none) */
    uint32_t has_weights; /* 1 if w came from the input, 0 if synthesised as 1 */
} gga_csr;

/* Reads a Galois .gr v1 file (replaces Graph::readFromGR, graph.h:599-682):
 * u64 version(=1), u64 sizeEdgeTy, u64 n, u64 m, u64 outIdx[n] (cumulative
 * ends), u32 outs[m], u32 pad if m is odd, u32 edgeData[m].  sizeEdgeTy must be
 * 0 (no weight section: weights are 1) or 4.  dst >= n is GGA_EFORMAT (the
 * reference only prints a warning, graph.h:669-670). */
int gga_csr_read_gr(const char *path, gga_csr *out);
/* Writes the same layout, always with a 4-byte weight section. */
int gga_csr_write_gr(const char *path, const gga_csr *g);
/* Builds a host CSR from .gr-style cumulative row ends (outIdx, graph.h:637,660-664). */
int gga_csr_from_row_end(uint32_t n, uint32_t m, const uint64_t *row_end,
                         const uint32_t *dst, const uint32_t *w, gga_csr *out);
void gga_csr_free(gga_csr *g);
/* Maximum out-degree and its first vertex (graph.h:338-358, first index on ties). */
uint32_t gga_csr_maxdeg(const gga_csr *g, uint32_t *first_index);

/* Seeded R-MAT generator (the reference used GTgraph, which it does not ship).
 * 2^scale vertices, edge_factor * 2^scale generated arcs, quadrant
 * probabilities a, b, c, d = 1 - a - b - c; weights uniform in [wlo, whi].
 * dedup != 0 removes duplicate arcs and self-loops. */
int gga_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint64_t seed, uint32_t wlo, uint32_t whi, int dedup, gga_csr *out);
/* rows x cols 4-neighbour grid (both arc directions), plus
 * round(shortcut_frac * undirected grid edges) random shortcut edges added in
 * both directions, no duplicates; one weight per undirected edge, uniform in
 * [wlo, whi]. */
int gga_gen_grid(uint32_t rows, uint32_t cols, double shortcut_frac, uint64_t seed,
                 uint32_t wlo, uint32_t whi, gga_csr *out);

/* ----------------------------------------------- device-resident graphs */
typedef struct gga_graph gga_graph;

/* Uploads a CSR given as .gr-style cumulative row ends (the survey-level
 * boundary: replaces Graph::cudaCopy + allocOnDevice, graph.h:534-553,683-699).
 * w == NULL means all weights 1.  The graph stays resident until gga_graph_free. */
int gga_graph_from_csr(uint32_t n, uint32_t m, const uint64_t *row_end,
                       const uint32_t *dst, const uint32_t *w, int device,
                       gga_graph **out);
/* Same, from a 0-based int32 CSR row_ptr[n + 1]. */
int gga_graph_from_csr32(uint32_t n, uint32_t m, const uint32_t *row_ptr,
                         const uint32_t *dst, const uint32_t *w, int device,
                         gga_graph **out);
/* readFromGR + cudaCopy in one call (bfs_sssp.cu:188,323). */
int gga_graph_load_gr(const char *path, int device, gga_graph **out);
int gga_graph_free(gga_graph *g);
uint32_t gga_graph_n(const gga_graph *g);
uint32_t gga_graph_m(const gga_graph *g);
/* Device bytes held by the graph (CSR + split + per-query workspace). */
uint64_t gga_graph_device_bytes(const gga_graph *g);

typedef struct gga_split_stats {
    uint32_t nnode;      /* myresult.nnode */
    uint32_t nedge;      /* myresult.nedge */
    uint32_t ns_nnode;   /* chunks = nodes after NodeSplit (graph.h:386-414) */
    uint32_t ns_nedge;   /* = nedge (splitting never adds edges) */
    uint32_t extranode;  /* ns_nnode - nnode */
    uint32_t maxdeg;     /* getMaxdeg, graph.h:338-347 */
    int32_t  level;      /* NodeSplitLevel L; -1 = histogram heuristic; 0 = explicit mda */
    uint32_t bucket;     /* histogram buckets when level < 0, else 0 (bfs_sssp.cu:177-180) */
    uint32_t mda;        /* MAX_EDGES_ALLOWED (the split threshold T) */
    uint32_t max_chunk;  /* largest chunk in edges (<= mda) */
    float    split_ms;   /* device time of the split pass */
    uint32_t exec_chunks;/* chunks the relax kernels walk: the split's chunks, with every chunk
                            longer than GGA_PIECE_MAX edges cut into pieces of at most that
                            many, so that a vertex the threshold leaves whole (L = 0 / 1, the
                            reference's dynamic-parallelism case) is still expanded by many
                            waves at once */
} gga_split_stats;
/* Longest chunk a wave expands: 4 rounds of 64 lanes x 4 edge slots. */
#define GGA_PIECE_MAX 1024u

/* Node splitting as a GPU prefix-sum pass (replaces Graph::NodeSplit,
 * graph.h:415-522, getSizeafterLoadBalance 386-414 and
 * getMaxAllowedOutDegree 359-385).  Threshold selection, in this order:
 *   mda > 0            : T = mda
 *   level > 0          : T = ceil(maxdeg / level)            (graph.h:390)
 *   level < 0          : T = histogram heuristic, `buckets` buckets (graph.h:359-385)
 * Every vertex of out-degree d > 0 becomes ceil(d / T) contiguous edge chunks
 * (first ones ceil(d / nchunks) edges, the last the remainder: INT_SPLIT,
 * graph.h:461-476); a vertex of degree 0 is one empty chunk.  Chunk ids equal
 * the reference's renumbered node ids.  Re-splitting replaces the old split. */
int gga_graph_split(gga_graph *g, uint32_t mda, int32_t level, uint32_t buckets,
                    gga_split_stats *st);
/* Copies the chunk layout to the host: chunk_begin[ns_nnode + 1] (0-based edge
 * offsets; the reference's psrc - 1) and owner[ns_nnode] (the reference's
 * parentNode).  Either pointer may be NULL. */
int gga_graph_chunks(gga_graph *g, uint32_t *chunk_begin, uint32_t *owner);

/* ------------------------------------------------------------ traversal */
#define GGA_MODE_TOPO        0  /* every chunk every iteration; host fixpoint loop (M_my_drelax + bfs_sssp.cu:372-468) */
#define GGA_MODE_WORKLIST    1  /* chunk worklist, wave64 ballot push; host-launched steps, optional hipGraph */
#define GGA_MODE_PERSISTENT  2  /* chunk worklist in one cooperative launch with a grid barrier per step */
#define GGA_MODE_AUTO        3  /* library choice from the degree distribution */

/* Instrumentation (the reference's metrics.h / my_drelax_metric per-block
 * clock spans, and convergence.h's per-iteration distance to the final
 * result).  One record per workgroup per relax step: start on the 100 MHz
 * device wall clock, duration in ticks, placement = XCC id << 8 | SE/SH/CU
 * bits of the HW_ID register. */
#ifndef GGA_WG_RECORD_DEFINED
#define GGA_WG_RECORD_DEFINED
typedef struct gga_wg_record {
    uint64_t start;
    uint32_t ticks;
    uint32_t place;
} gga_wg_record;
#endif
/* One row per relax step: vertices whose distance still differs from the
 * final one, and the L2 norm of the difference computed as printConv_rtime
 * does (convergence.h:100-140: squares of unsigned differences, mod 2^32). */
typedef struct gga_conv_row {
    uint32_t step;
    uint32_t diff_count;
    double   norm2;
} gga_conv_row;

typedef struct gga_opts {
    int32_t  mode;        /* GGA_MODE_* (default GGA_MODE_AUTO) */
    uint32_t inf;         /* unreached value: GGA_INF (0 -> default) or GGA_REF_INF */
    uint32_t max_iters;   /* 0 -> n + 1 (the serial bound, graph.h:122) */
    uint32_t block;       /* WORKLIST: workgroup size, 128, 256 or 512 threads (0 -> the tuned
                             size, else 256; GGA_EINVAL otherwise); PERSISTENT runs 256 */
    uint32_t sync_every;  /* WORKLIST: steps launched per host convergence check (0 -> 8, the
                             first check after one step more than the graph's last query
                             needed, at most 32) */
    uint32_t use_graph;   /* multi-source engines: replay steps from captured hipGraphs (six
                             steps per graph) */
    uint32_t count;       /* 1 -> collect edges_inspected / chunks_expanded / relaxations, and
                             time each relax / list launch (ktime_ms, ctime_ms) */
    uint32_t grid;        /* PERSISTENT / WORKLIST grid in workgroups, 0 -> occupancy bound */
    uint32_t dense_min;   /* WORKLIST / PERSISTENT: a step whose list holds >= dense_min chunks
                             records the next frontier as step stamps + compaction instead of
                             queue pushes; 0 -> max(4096, n / 64); 0xFFFFFFFF -> never */
    uint32_t pull_alpha;  /* gga_multi_bfs: a level runs bottom-up (pull over in-edges) when the
                             frontier's out-edges x pull_alpha > m; 0 -> 14; 0xFFFFFFFF -> never */
    uint32_t ms_layout;   /* gga_multi_sp: GGA_MS_* engine layout (0 -> GGA_MS_AUTO) */
    uint32_t delta;       /* SSSP, WORKLIST / PERSISTENT: near-far bucket width -- an improved vertex is listed
                             only while its distance is below the current threshold, the rest
                             wait until the threshold passes them (Davidson et al. 2014), so
                             fewer vertices are expanded more than once; 0 -> the tuned width
                             (gga_tune), else 32 mean edge weights on mesh-like graphs (max
                             degree within 4x the mean + 8), else off; 0xFFFFFFFF -> off */
    /* WORKLIST instrumentation (both off when NULL) */
    gga_wg_record *timeline;    /* host buffer: grid records per step, steps in order */
    uint32_t timeline_cap;      /* records it holds; steps beyond it run uninstrumented */
    uint32_t conv_cap;          /* rows conv_rows holds */
    const uint32_t *conv_final; /* host: the final distances (n), e.g. from a previous run */
    gga_conv_row *conv_rows;    /* host: one row per relax step */
} gga_opts;
/* Fills defaults (all zero = defaults). */
void gga_opts_init(gga_opts *o);

typedef struct gga_stats {
    uint32_t iterations;      /* relax steps run (reference `iterations`) */
    uint32_t reached;         /* vertices with dist < inf */
    float    time_ms;         /* device time, dist init -> convergence (reference `time`, bfs_sssp.cu:347-469) */
    float    ktime_ms;        /* relax-kernel device time (reference `ktime`); measured with
                                 per-launch events, so only when opts.count is set (they cost
                                 a few us of dispatch gap each) or in PERSISTENT mode */
    float    sktime_ms;       /* child-sync time: always 0 (no child copies; reference `sktime`) */
    float    ctime_ms;        /* WORKLIST: list-compaction kernel time (after flag steps) */
    int32_t  mode;            /* mode actually used */
    uint64_t edges_inspected; /* I_e  (opts.count) */
    uint64_t chunks_expanded; /* I_v  (opts.count) */
    uint64_t relaxations;     /* R: successful atomicMin (opts.count) */
    uint64_t m_reach;         /* sum of out-degrees of reached vertices */
    uint64_t launches;        /* relax-kernel launches */
    uint32_t pull_levels;     /* gga_multi_bfs: levels run bottom-up */
    uint32_t timeline_records;/* records written to opts.timeline */
    uint32_t timeline_grid;   /* workgroups per step in it */
    uint32_t conv_steps;      /* rows written to opts.conv_rows */
} gga_stats;

/* BFS levels from src (replaces the na_bfs path: dist init bfs_sssp.cu:302-321,
 * fixpoint loop 372-468, M_my_drelax cudaKernels.h:285-344).  out: n uint32 on
 * the host, or NULL to keep the result in the graph's device buffer. */
int gga_bfs(gga_graph *g, uint32_t src, uint32_t *out, const gga_opts *o, gga_stats *st);
/* SSSP distances from src with the stored uint32 weights (the na_sssp path). */
int gga_sssp(gga_graph *g, uint32_t src, uint32_t *out, const gga_opts *o, gga_stats *st);
/* SSSP that allocates nothing (no hipMalloc / hipFree / event or stream
 * creation).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory once set up -- tight serving loops, harnesses that
 * forbid allocation after initialisation.  The graph must be uploaded and
 * split beforehand (GGA_ENOSPLIT otherwise); d_out is the caller's device
 * buffer of n uint32 and receives the distances.  The query's workspace
 * (step stamps, flag bytes, chunk lists, control words, counters, near-far
 * arrays) lives in one static device array of GGA_SCRATCH_BYTES (128 MiB)
 * per device, held from module load to process end; a graph whose workspace
 * exceeds it -- 13 n + 8 C bytes plus ~300 KB, + 4 n with near-far -- is
 * GGA_ERANGE.  Calls share that array: the library serialises them
 * (one at a time per process), and each call starts its contents over.
 * Instrumentation (opts.timeline / conv_rows) is GGA_EINVAL; ktime_ms and
 * ctime_ms stay 0 outside PERSISTENT mode. */
int gga_sssp_noalloc(gga_graph *g, uint32_t src, uint32_t *d_out, const gga_opts *o, gga_stats *st);
/* Copies the last result (n uint32) to the host (bfs_sssp.cu:489). */
int gga_graph_copy_dist(gga_graph *g, uint32_t *out);
/* Device pointer of the last result (n uint32), valid until the next query. */
const uint32_t *gga_graph_dist_device(gga_graph *g);

/* --------------------------------------------------- checks and outputs */
typedef struct gga_verify_result {
    uint64_t nerr;        /* edges with w > 0 and dist[u] + w < dist[v] (verifysolution, myutils.h:1-18) */
    uint64_t untight;     /* reached v != src with no u: dist[u] + w == dist[v] */
    uint32_t src_ok;      /* dist[src] == 0 */
    uint32_t reached;     /* vertices with dist < inf */
} gga_verify_result;
/* Host-side fixpoint + tightness check of dist against the CSR.  nerr == 0,
 * untight == 0 and src_ok == 1 together fix the shortest-path distances when
 * every weight is >= 1.  w == NULL means BFS (weights 1). */
int gga_verify(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst, const uint32_t *w,
               uint32_t src, const uint32_t *dist, uint32_t inf, gga_verify_result *r);
/* Same check on the device against the graph's last result (for sizes where
 * a host copy is wasteful). */
int gga_graph_verify(gga_graph *g, uint32_t src, int bfs, uint32_t inf, gga_verify_result *r);

/* List-scheduling estimate of the relax time from a timeline (combineClocks,
 * metrics.h:70-92): each step's workgroup durations, in launch order, go to
 * the least-loaded of `slots` execution slots (CUs x resident workgroups);
 * returns the sum over steps of the busiest slot, in ms (100 MHz ticks). */
double gga_timeline_makespan(const gga_wg_record *r, uint32_t nrec, uint32_t grid, uint32_t slots);
/* Appends the convergence block of printConv_rtime_init / _rtime / _final
 * (convergence.h:81-140) to path: header, "Iteration:<step> = <norm2>(<diff>)"
 * per row, "Total Iteration : <n>     Max Dist : <max_dist>". */
int gga_write_conv(const char *path, int32_t level, uint32_t maxdeg, const char *graph, const char *algo,
                   const gga_conv_row *rows, uint32_t nrows, uint32_t max_dist);

/* gga_checksum of the last result, computed on the device. */
int gga_graph_checksum(gga_graph *g, uint64_t *sum);

/* Degree distribution file (printfDegreeDistribution, graph.h:98-114, written
 * as DD_<graph>.txt when bfs_sssp.cu:196-200 is enabled): one "deg,count"
 * line per out-degree that occurs, ascending, from the host CSR row_ptr[n + 1]. */
int gga_write_degree_dist(const char *path, uint32_t n, const uint32_t *row_ptr);
/* Distance file: one "%u\n" per original vertex (Ary2File, myutils.h:51-62). */
int gga_write_dist(const char *path, const uint32_t *dist, uint32_t n);

typedef struct gga_result_row {
    char exename[100];
    char algo[10];
    char graphfile[200];
    uint32_t bucket, nnode, nedge, ns_nnode, ns_nedge, extranode, maxdeg;
    int32_t  level;
    uint32_t mda, iterations, runtime_ms;
    float time_ms, ktime_ms, sktime_ms;
} gga_result_row;
/* Appends the 17-field CSV row of writeResult (myutils.h:86-113) to path. */
int gga_write_result_row(const char *path, const gga_result_row *r);
/* Graph name = last path component (getGraphname, myutils.h:76-84). */
void gga_graph_name(const char *path, char *out, size_t cap);

/* ------------------------------------------- multi-source engine (32 lanes)
 * No reference counterpart (the reference runs one source per process,
 * bfs_sssp.cu:203).  32 independent BFS/SSSP traversals ("lanes") share every
 * relax step of one graph: distances are stored vertex-major (the 32 lanes of
 * a vertex are one 128-byte line), a step expands each listed chunk only for
 * the lanes whose distance at its owner dropped in the previous step, and each
 * lane converges on its own.  A finished lane can be restarted with a new
 * source while the others keep running (the serving loop's continuous
 * batching).  Results per lane equal gga_bfs / gga_sssp from that source. */
#define GGA_MS_LANES 32
typedef struct gga_ms gga_ms;
/* Engine on a resident graph (split with the defaults first if it is not),
 * with a stream of its own.  Device memory: 136 B per vertex + 8 B per chunk
 * + the per-tile scratch.
 * inf: unreached value (0 -> GGA_INF).  layout: GGA_MS_VERTEX_MAJOR (the 32
 * lanes of a vertex in one 128-byte line: lanes that meet at the same
 * vertices share every gather -- small-world graphs), GGA_MS_LANE_MAJOR (each
 * lane's distances contiguous: lanes whose frontiers rarely overlap, e.g.
 * meshes, gather only their own 4 bytes), or GGA_MS_AUTO (the tuner's choice
 * for the graph, gga_tune_layout; untuned: lane-major on mesh-like graphs,
 * else vertex-major).  Re-splitting the graph invalidates it. */
#define GGA_MS_AUTO          0
#define GGA_MS_VERTEX_MAJOR  1
#define GGA_MS_LANE_MAJOR    2
int gga_ms_create(gga_graph *g, uint32_t inf, uint32_t layout, gga_ms **out);
int gga_ms_free(gga_ms *e);
/* Near-far buckets per SSSP lane (as opts.delta for one source): a lane's
 * improvement at or beyond its threshold waits until a refill reaches it.
 * Set at creation from the graph (gga_tune's width, or 16 mean weights on
 * mesh-like graphs -- half the single-source default -- else off); delta 0 or
 * 0xFFFFFFFF turns it off.  gga_multi_sp uses the same default.  Only while
 * no lane runs; near-far engines run in step mode (GGA_MODE_WORKLIST).
 * +4 B per vertex (the far masks). */
int gga_ms_set_delta(gga_ms *e, uint32_t delta);
/* Device bytes gga_ms_create would take on this graph (near-far included). */
uint64_t gga_ms_graph_bytes(gga_graph *g);
/* Restarts the lanes in `lanes` (bit i = lane i) at src[i] (src holds 32
 * entries; only those of `lanes` are read).  Lanes in unit_lanes count hops
 * (BFS levels), the others add edge weights.  A lane still running is EINVAL. */
int gga_ms_start(gga_ms *e, uint32_t lanes, const uint32_t *src, uint32_t unit_lanes);
/* Runs up to max_steps steps (0: until every lane converged).  mode:
 * GGA_MODE_WORKLIST (three kernels per step, three more with near-far lanes,
 * `sync_every` steps per host check, replayed from captured hipGraphs when
 * use_graph) or GGA_MODE_PERSISTENT (one cooperative launch).  o may be NULL.  *live = lanes with work left; a lane
 * started and not live has finished. */
int gga_ms_run(gga_ms *e, uint32_t max_steps, const gga_opts *o, uint32_t *live);
/* Asynchronous form of a run of nsteps steps, for callers that drive several
 * engines at once: launch enqueues the steps and a read-back on the engine's
 * own stream and returns; collect waits for them and reports *live as
 * gga_ms_run does.  Every launch is followed by one collect. */
int gga_ms_launch(gga_ms *e, uint32_t nsteps, const gga_opts *o);
int gga_ms_collect(gga_ms *e, uint32_t *live);
/* 1 when a launch has completed (collect will not wait) or none is pending,
 * 0 while it runs, < 0 on error. */
int gga_ms_ready(gga_ms *e);
/* Lanes started and finished since their last start (not yet restarted). */
uint32_t gga_ms_done(const gga_ms *e);
/* Fingerprints of finished lanes' distances: sums[i] = gga_checksum of lane
 * i's n values for i in lanes (0 elsewhere; sums holds 32). */
int gga_ms_checksum(gga_ms *e, uint32_t lanes, uint64_t *sums);
/* Device bytes of an engine on a graph with n vertices and `chunks` chunks. */
uint64_t gga_ms_bytes(uint32_t n, uint32_t chunks);
/* sum over v of splitmix64(v << 32 | dist[v]) mod 2^64: an order-independent
 * fingerprint used to compare results without moving them. */
uint64_t gga_checksum(const uint32_t *dist, uint32_t n);
/* Copies the distances of finished lanes: out[i] (n uint32) for each lane i in
 * `lanes`; host pointers, or device pointers when on_device. */
int gga_ms_fetch(gga_ms *e, uint32_t lanes, uint32_t *const *out, int on_device);
/* Engine totals since creation: iterations = steps, edges_inspected =
 * (edge, lane) relax checks, relaxations, ktime_ms = step kernel time. */
int gga_ms_stats(const gga_ms *e, gga_stats *st);

/* k <= 32 sources in one traversal (config: batched multi-source SSSP / BFS):
 * out = k x n row-major (row i = source srcs[i]) or NULL to keep the results on
 * the device; lanes in unit_lanes are BFS.  o->mode GGA_MODE_WORKLIST or
 * GGA_MODE_PERSISTENT (AUTO -> WORKLIST).  st: iterations = steps, time_ms =
 * device time start -> convergence, reached / m_reach summed over sources. */
int gga_multi_sp(gga_graph *g, const uint32_t *srcs, uint32_t k, uint32_t unit_lanes, uint32_t *out,
                 const gga_opts *o, gga_stats *st);

/* Multi-source BFS with bitmask lanes: k <= 256 sources in one cooperative
 * launch (per vertex 4 x 64-bit visited / frontier / next words; levels pushed
 * top-down from small frontiers and pulled bottom-up over in-edges when the
 * frontier is large, see opts.pull_alpha).  out = k x n uint32 levels
 * row-major (row i = source srcs[i]), unreached = o->inf; NULL keeps them on
 * the device.  Builds and caches the transposed graph on first use.
 * st: iterations = levels, reached (clamped to uint32) / m_reach summed over
 * sources, time_ms = device time init -> all levels written. */
#define GGA_MSBFS_LANES 256
int gga_multi_bfs(gga_graph *g, const uint32_t *srcs, uint32_t k, uint32_t *out, const gga_opts *o,
                  gga_stats *st);

/* ------------------------------------------------------------ autotuner
 * Replaces the histogram guess getMaxAllowedOutDegree (graph.h:359-385) and
 * the NodeSplitLevel sweep of mysc (mysc:20-51): the histogram threshold
 * (10 buckets, bfs_sssp.cu:166) seeds the candidates {seed, 16, 32, 64, 128,
 * 256, 1024, maxdeg}; each is timed on one query from src (best of 2 device
 * times), the fastest split is kept, then the worklist grid (default, 1/2,
 * 1/4), the push / compaction switch (dense_min x 1/4, x 4) and the launch
 * mode with the near-far width (worklist / persistent, off / 2 / 8 / 32 mean
 * weights) are timed the same way.  The graph keeps the choices: later
 * queries whose opts leave mode AUTO and grid, dense_min, delta at 0 use
 * them.  GGA_TUNE_LAYOUT also times both multi-source engine layouts on 32
 * spread sources.  Results never depend on any choice. */
#define GGA_TUNE_BFS     1u   /* tune with BFS queries (default SSSP) */
#define GGA_TUNE_LAYOUT  2u   /* also measure the multi-source layout */
#define GGA_TUNE_MULTI   4u   /* objective: a 32-lane gga_multi_sp run from 32 spread sources
                                 (split threshold only; grid and switch are left default) */
#define GGA_TUNE_MAX     16
typedef struct gga_tune_result {
    uint32_t mda;             /* chosen split threshold (the graph is left split with it) */
    uint32_t grid;            /* chosen worklist grid (workgroups of `block` threads) */
    uint32_t dense_min;       /* chosen push / compaction switch */
    uint32_t layout;          /* GGA_MS_VERTEX_MAJOR or GGA_MS_LANE_MAJOR */
    int32_t  mode;            /* chosen GGA_MODE_WORKLIST / GGA_MODE_PERSISTENT (what AUTO then uses) */
    uint32_t delta;           /* chosen near-far width (0xFFFFFFFF: plain Bellman-Ford) */
    uint32_t seed_mda;        /* the histogram heuristic's threshold */
    uint32_t ntried;          /* split thresholds timed */
    float    seed_ms, best_ms;
    float    layout_ms[2];    /* vertex-major, lane-major (GGA_TUNE_LAYOUT) */
    uint32_t tried_mda[GGA_TUNE_MAX];
    float    tried_ms[GGA_TUNE_MAX];
    uint32_t block;           /* chosen worklist workgroup size (128, 256 or 512 threads) */
    float    block_ms[3];     /* best query time at 128, 256, 512 threads (single-source objective) */
} gga_tune_result;
int gga_tune(gga_graph *g, uint32_t src, uint32_t flags, gga_tune_result *r);
/* Engine layout for the graph: the measured one after gga_tune with
 * GGA_TUNE_LAYOUT, else vertex-major. */
uint32_t gga_tune_layout(gga_graph *g);

/* ------------------------------------------------------------- serving
 * No reference counterpart (one query per process in the reference).  A
 * server keeps graphs resident in HBM and answers BFS/SSSP queries on them:
 * each graph gets, when a query needs it, a 32-lane multi-source engine (up
 * to engines_per_graph of them) whose device memory (gga_ms_bytes) is
 * admitted against hbm_budget, evicting idle engines least-recently-used
 * first; a query takes one lane from admission to convergence, and lanes
 * freed by finished queries are refilled from the waiting queries at the next
 * round (continuous batching).  Every round launches `steps_per_round` steps
 * on every engine with running lanes (one stream per engine, so engines
 * overlap on the device).  Results do not depend on the mix: each lane is the
 * single-source computation. */
typedef struct gga_server gga_server;
typedef struct gga_server_opts {
    uint64_t hbm_budget;      /* bytes of engine memory (0 -> 8 GiB) */
    uint32_t steps_per_round; /* 0 -> 12 (two captured six-step graphs per round) */
    uint32_t lanes;           /* lanes used per engine, 1..32 (0 -> 32) */
    uint32_t use_graph;       /* steps replayed from captured hipGraphs */
    uint32_t dense_min;       /* engines' push / compaction switch (0 -> default) */
    uint32_t host_results;    /* queries may ask for host copies: staging counted in the budget */
    uint32_t dry_run;         /* no device: engines are only accounted and a query
                                 completes dry_steps rounds after admission (tests of
                                 admission and batching on hosts without a GPU) */
    uint32_t dry_steps;       /* 0 -> 3 */
    uint32_t layout;          /* GGA_MS_* layout of every engine (0 -> the graph's tuned layout,
                                 else lane-major: fewer bytes per step when an engine runs a few
                                 lanes; GGA_MS_VERTEX_MAJOR serves full 32-lane bursts faster) */
    uint32_t engines_per_graph; /* engines one graph may hold at once, 1..GGA_SERVER_MAX_ENGINES
                                   (0 -> 1): a further engine takes a graph's queries while the
                                   others' lanes are all busy */
    int32_t  mesh_reserve;    /* eighths of each CU's workgroup slots that engines on graphs
                                 that are not mesh-like leave free (their relax grid is cut by
                                 as much), so that the short launches of engines on mesh-like
                                 graphs -- thousands of steps per query -- do not queue behind
                                 theirs: 0 -> 2, 1..6 as given, -1 -> none */
} gga_server_opts;
#define GGA_SERVER_MAX_ENGINES 4
typedef struct gga_query {
    uint32_t graph;           /* id from gga_server_add_graph */
    uint32_t src;
    uint32_t bfs;             /* 1: BFS levels, 0: SSSP */
    uint32_t out_on_device;   /* out is a device pointer */
    double   arrival;         /* server clock (s); <= 0: now.  Replays: offset from the start */
    uint32_t *out;            /* n uint32 for the result, or NULL for its fingerprint only */
} gga_query;
typedef struct gga_query_result {
    uint64_t id;
    uint32_t graph, src, bfs, lane;
    int32_t  status;          /* GGA_OK; GGA_EBUDGET if its engine exceeds the whole budget */
    uint32_t engine;          /* which of its graph's engines ran it */
    double   arrival, admitted, done;   /* seconds (replays: from the replay's start) */
    uint64_t checksum;        /* gga_checksum of the distances */
} gga_query_result;
typedef struct gga_server_counters {
    uint64_t submitted, completed, rejected, rounds, evictions, engines_created;
    uint64_t bytes_in_use, peak_bytes;
    uint32_t engines_resident;
    uint32_t peak_running;    /* most queries running at once */
    uint64_t joined;          /* queries admitted into an engine while other lanes of it were
                                 running (continuous batching: work of a new query overlaps
                                 work already on the device) */
} gga_server_counters;

/* Threading: gga_server_submit, gga_server_poll and gga_server_now may be
 * called from any thread while one thread drives pump / drain / replay;
 * add_graph, stats and free run on that thread (or with no other call in
 * flight). */
int gga_server_create(const gga_server_opts *o, gga_server **out);
int gga_server_free(gga_server *s);
/* The graph stays owned by the caller and must outlive the server. */
int gga_server_add_graph(gga_server *s, gga_graph *g, uint32_t *id);
/* Dry-run servers: a graph known only by its size (n vertices, chunks). */
int gga_server_add_graph_dry(gga_server *s, uint32_t n, uint32_t chunks, uint32_t *id);
/* Seconds since the server was created. */
double gga_server_now(const gga_server *s);
/* Queues a query (thread-safe); *id identifies its result. */
int gga_server_submit(gga_server *s, const gga_query *q, uint64_t *id);
/* One round: admission, lane restarts, steps, result delivery.  *active =
 * lanes still running. */
int gga_server_pump(gga_server *s, uint32_t *active);
/* Up to cap finished results, oldest first. */
int gga_server_poll(gga_server *s, gga_query_result *res, uint32_t cap, uint32_t *count);
/* Pumps until no query waits or runs. */
int gga_server_drain(gga_server *s);
/* Open-loop replay: trace[i] is submitted when the clock reaches its arrival
 * offset (trace sorted by arrival), rounds run meanwhile; res[i] is trace[i]'s
 * result. */
int gga_server_replay(gga_server *s, const gga_query *trace, uint32_t nq, gga_query_result *res);
/* The same trace answered one query at a time with gga_bfs / gga_sssp in
 * `mode` (GGA_MODE_PERSISTENT: one launch per query) -- the baseline the
 * batched server is measured against. */
int gga_server_replay_plain(gga_server *s, const gga_query *trace, uint32_t nq, int32_t mode,
                            gga_query_result *res);
int gga_server_stats(const gga_server *s, gga_server_counters *st);

#ifdef __cplusplus
}
#endif
#endif /* GGA_H */
