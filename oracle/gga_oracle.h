/*
 * gga_oracle.h -- CPU oracle for the BFS/SSSP node-splitting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product library never links or calls it.
 *
 * Each orc_* function restates one reference function
 * (Programs/CUDA_src/Kepler_BFS_SSSP/...) in plain C, on a 0-based CSR
 * row_ptr[n + 1] / dst[m] / w[m] (w == NULL: BFS, every weight 1).
 *
 * How each function is pinned (the tests that would catch a wrong sign,
 * index or dropped term):
 *   orc_serial_sp      golden vectors written by scipy.sparse.csgraph (dijkstra,
 *                      breadth_first_order) and networkx; brute-force
 *                      Floyd-Warshall on tiny graphs; fixpoint + tightness.
 *   orc_verify         planted errors in otherwise-correct distance arrays.
 *   orc_tight          planted too-small distances (a fixpoint is not enough).
 *   orc_maxdeg         numpy argmax (first index on ties).
 *   orc_mda_level,
 *   orc_mda_histogram,
 *   orc_split_count    the reference's published result.vi rows (MAX_EDGES_ALLOWED
 *                      and ns_nnode for 704 runs) recomputed from its published
 *                      degree histograms DD_rmat20.gr.txt / DD_webs.gr.txt.
 *   orc_node_split     closed forms: chunk sizes sum to the degree, first
 *                      chunks ceil(d / k) edges, count == orc_split_count.
 *   orc_split_relax    equality with orc_serial_sp (the split graph with
 *                      pseudo-children must give the same distances).
 *   orc_dijkstra,
 *   orc_bfs_queue      textbook routines (fast enough for the full-size
 *                      configs), pinned by equality with orc_serial_sp and the
 *                      scipy golden vectors.
 */
#ifndef GGA_ORACLE_H
#define GGA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Seidel Bellman-Ford of Graph::findSerialSSSP / findSerialBFS
 * (graph.h:115-184): at most n sweeps over vertices in id order and their
 * out-edges in CSR order, stop after a sweep with no update.  `inf` is the
 * unreached value (the reference's MYINFINITY = 1e9; pass 0xFFFFFFFF for the
 * uint32 convention).  A relaxation whose sum would reach or pass `inf`, or wrap,
 * is skipped (with inf = 1e9 and weights < 3.29e9 this is exactly the
 * reference's comparison).  Returns the number of sweeps; *max_dist receives
 * the largest finite distance (max_sssp / max_bfs). */
uint32_t orc_serial_sp(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                       const uint32_t *w, uint32_t src, uint32_t inf,
                       uint32_t *dist, uint32_t *max_dist);

/* verifysolution (myutils.h:1-18): number of edges with w > 0 and
 * dist[u] + w < dist[v]; sources at `inf` are skipped (no wrap). */
uint64_t orc_verify(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                    const uint32_t *w, const uint32_t *dist, uint32_t inf);

/* Number of reached v != src without a tight in-edge (dist[u] + w == dist[v]);
 * adds 1 when dist[src] != 0. */
uint64_t orc_tight(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                   const uint32_t *w, uint32_t src, const uint32_t *dist, uint32_t inf);

/* getMaxdeg / getMaxdegIndex (graph.h:338-358). */
uint32_t orc_maxdeg(uint32_t n, const uint32_t *row_ptr, uint32_t *first_index);

/* MAX_EDGES_ALLOWED = (maxdeg + L - 1) / L in unsigned arithmetic (graph.h:390). */
uint32_t orc_mda_level(uint32_t maxdeg, int32_t level);

/* getMaxAllowedOutDegree (graph.h:359-385) over a degree array: float32 bucket
 * index (int)((deg / (float)(maxdeg + 1)) * buckets), most populated bucket
 * (first on ties), result ((maxdeg + 1) / buckets) * (idx + 1). */
uint32_t orc_mda_histogram(uint32_t n, const uint32_t *deg, int32_t buckets);

/* Same, from a (degree, count) histogram as in DD_*.gr.txt (graph.h:98-114). */
uint32_t orc_mda_histogram_dd(uint32_t nrows, const uint32_t *deg, const uint32_t *cnt,
                              int32_t buckets);

/* getSizeafterLoadBalance node count (graph.h:399-412):
 * sum over v of (deg == 0 ? 1 : ceil(deg / mda)). */
uint64_t orc_split_count(uint32_t n, const uint32_t *deg, uint32_t mda);

/* NodeSplit layout with INT_SPLIT (graph.h:415-522, bfs_sssp.cu:60):
 * chunk_begin[C + 1] = 0-based edge offsets (the reference's psrc - 1),
 * owner[C] = original vertex (the reference's parentNode), child_count[C]
 * (non-zero only at a parent), presum[n + 1] (new parent id of v = v + presum[v]).
 * Any output may be NULL.  Returns C. */
uint32_t orc_node_split(uint32_t n, const uint32_t *row_ptr, uint32_t mda,
                        uint32_t *chunk_begin, uint32_t *owner,
                        uint32_t *child_count, uint32_t *presum);

/* The reference's device algorithm run sequentially: build the split graph
 * (renumbered nodes, destinations remapped to parent ids, graph.h:500-502),
 * initialise the source parent and its children to 0 (bfs_sssp.cu:302-318),
 * then sweep M_my_drelax (cudaKernels.h:285-344: relax, and atomicMin on every
 * child of the destination) until no change.  dist receives the parents'
 * values in original vertex order (Ary2File's IsParent filter, myutils.h:51-62).
 * Returns the number of sweeps. */
uint32_t orc_split_relax(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                         const uint32_t *w, uint32_t mda, uint32_t src, uint32_t inf,
                         uint32_t *dist);

/* Binary-heap Dijkstra (textbook; not a reference function). */
void orc_dijkstra(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                  const uint32_t *w, uint32_t src, uint32_t inf, uint32_t *dist);
/* FIFO-queue BFS levels (textbook; not a reference function). */
void orc_bfs_queue(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst,
                   uint32_t src, uint32_t inf, uint32_t *dist);

/* Sum of out-degrees of vertices with dist < inf (m_reach for GTEPS). */
uint64_t orc_m_reach(uint32_t n, const uint32_t *row_ptr, const uint32_t *dist, uint32_t inf);

#ifdef __cplusplus
}
#endif
#endif
