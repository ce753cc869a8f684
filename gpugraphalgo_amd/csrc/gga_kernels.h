/*
 * gga_kernels.h -- thin C-ABI between the C host layer and the HIP kernels.
 * Every launcher enqueues on `s` and returns the hipError_t of the launch.
 */
#ifndef GGA_KERNELS_H
#define GGA_KERNELS_H
#include <stdint.h>
#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#include "gga_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- split pass (split.hip) ---- */
/* key = (maxdeg << 32) | ~first_index, atomically maxed into *d_key (zero it first). */
int ggak_degree_max(const uint32_t *row, uint32_t n, uint64_t *d_key, hipStream_t s);
/* float32 bucket histogram of getMaxAllowedOutDegree (graph.h:359-385); buckets <= 4096. */
int ggak_degree_hist(const uint32_t *row, uint32_t n, uint32_t maxdeg, uint32_t buckets,
                     uint32_t *d_hist, hipStream_t s);
/* Scratch words ggak_split_scan needs. */
uint32_t ggak_split_tmp_words(uint32_t n);
/* cstart[v] = exclusive prefix of max(1, ceil(deg / mda)); cstart[n] = C. */
int ggak_split_scan(const uint32_t *row, uint32_t n, uint32_t mda, uint32_t *cstart,
                    uint32_t *d_tmp, hipStream_t s);
/* The same with `empty` (0 or 1) chunks for a degree-0 vertex. */
int ggak_split_scan_ex(const uint32_t *row, uint32_t n, uint32_t mda, uint32_t empty, uint32_t *cstart,
                       uint32_t *d_tmp, hipStream_t s);
/* Chunk descriptors cbeg[C + 1], cown[C]; d_maxchunk[GGA_MAXCHUNK_SLOTS] (zeroed)
 * receives per-slot maxima of the chunk sizes. */
#define GGA_MAXCHUNK_SLOTS 64u
int ggak_split_fill(const uint32_t *row, const uint32_t *cstart, uint32_t n, uint32_t m,
                    uint32_t *cbeg, uint32_t *cown, uint32_t *d_maxchunk, hipStream_t s);

/* out[len + 1] = exclusive scan of in[len] (out must not alias in); d_tmp needs
 * ggak_split_tmp_words(len) words. */
int ggak_scan_u32(const uint32_t *in, uint32_t len, uint32_t *out, uint32_t *d_tmp, hipStream_t s);

/* ---- static workspace of the allocation-free query (relax.hip) ---- */
#define GGA_SCRATCH_BYTES (128ull << 20)
/* *ptr = this device's address of the static scratch array (looked up once
 * per device with hipGetSymbolAddress, that device current); returns 0 or a
 * hipError_t. */
int ggak_scratch(int device, void **ptr);

/* ---- traversal (relax.hip) ---- */
int ggak_fill_u32(uint32_t *p, uint32_t value, uint64_t count, hipStream_t s);
int ggak_query_init(gga_dev_graph g, gga_dev_work w, uint32_t src, uint32_t epoch, hipStream_t s);
int ggak_topo_step(gga_dev_graph g, gga_dev_work w, int bfs, int count, hipStream_t s);
/* tl != NULL: instrumented (counting) variant; workgroup b writes tl[b].
 * block: workgroup size, 128, 256 or 512 threads. */
int ggak_wl_step(gga_dev_graph g, gga_dev_work w, uint32_t phase, int bfs, int count,
                 uint32_t grid, uint32_t block, gga_wg_record *tl, hipStream_t s);
/* rows[0] += #v with dist != fin, rows[1] += sum of (fin - dist)^2 in uint32. */
int ggak_conv(const uint32_t *dist, const uint32_t *fin, uint32_t n, uint64_t *rows, hipStream_t s);
/* After ggak_wl_step(phase): builds the next list from step stamps when that
 * step was a flag step (its list held >= w.dense_min chunks); else a no-op. */
int ggak_compact(gga_dev_graph g, gga_dev_work w, uint32_t phase, hipStream_t s);
/* Near-far (w.last != NULL): after ggak_compact(phase), when the next list is
 * empty, lists the deferred vertices under the next threshold. */
int ggak_refill(gga_dev_graph g, gga_dev_work w, uint32_t phase, hipStream_t s);
/* Occupancy-bound grid (workgroups of `block` threads) for the worklist
 * kernel, or for the persistent kernel (always 256 threads); 0 on error. */
uint32_t ggak_wl_grid(int device, int persistent, int bfs, int count, uint32_t block);
/* One cooperative launch running worklist steps until the list is empty or
 * max_steps; grid must not exceed ggak_wl_grid(dev, 1, ...). */
int ggak_persistent(gga_dev_graph g, gga_dev_work w, int bfs, int count, uint32_t grid,
                    uint32_t max_steps, uint32_t spin_limit_ms, hipStream_t s);
/* reached count -> ctl[RED] (u64), m_reach -> ctl[RED + 2] (u64). */
int ggak_reach(gga_dev_graph g, gga_dev_work w, hipStream_t s);
/* fingerprint of dist (gga_checksum) -> ctl[RED + 4] (u64). */
int ggak_checksum(gga_dev_graph g, gga_dev_work w, hipStream_t s);
/* nerr -> ctl[RED] (u64), untight -> ctl[RED + 2] (u64); d_tight: n bytes scratch. */
int ggak_verify(gga_dev_graph g, gga_dev_work w, uint32_t src, int bfs, uint8_t *d_tight,
                hipStream_t s);

/* ---- multi-source engine (ms.hip), 32 lanes; `step` is the engine's step number */
int ggak_ms_clear(gga_dev_ms ms, hipStream_t s);
/* Restart `lanes` at src.v[lane] (unit: which of them count hops) and list
 * them for step `step`. */
int ggak_ms_start(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, uint32_t unit, gga_ms_sources src,
                  uint32_t step, hipStream_t s);
/* One step: relax + list rebuild; ctl[TAIL((step+1)%3)] / ctl[LANES((step+1)%3)] give
 * the next list length and its live lanes. */
int ggak_ms_step(gga_dev_graph g, gga_dev_ms ms, uint32_t step, int count, uint32_t grid, hipStream_t s);
uint32_t ggak_ms_grid(int device, int persistent, int count);
/* Steps step0, step0 + 1, ... in one cooperative launch until the list is empty
 * or max_steps ran; ctl[STEPS] grows by the number of steps run. */
int ggak_ms_persistent(gga_dev_graph g, gga_dev_ms ms, uint32_t step0, uint32_t max_steps, int count,
                       uint32_t grid, uint32_t spin_limit_ms, hipStream_t s);
/* out.p[lane][v] = D[v][lane] for the lanes in `lanes` (device pointers). */
int ggak_ms_harvest(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, gga_ms_outputs out, hipStream_t s);
/* Reached (vertex, lane) pairs of `lanes` and their out-degree sum -> cnt word 3
 * partials; ggak_ms_fold then leaves them in ctl[RED], ctl[RED + 2] (uint64). */
int ggak_ms_reach(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, hipStream_t s);
/* sums[lane] += fingerprint of lane's distances (sums zeroed first; 32 uint64). */
int ggak_ms_checksum(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, uint64_t *sums, hipStream_t s);
/* Striped work counters -> ctl[CNT] (3 x uint64), reach partials -> ctl[RED]. */
int ggak_ms_fold(gga_dev_ms ms, hipStream_t s);

/* ---- multi-source BFS with bitmask lanes (msbfs.hip) */
/* indeg[col[e]] += 1 for every edge (indeg zeroed first). */
int ggak_indeg(const uint32_t *col, uint32_t m, uint32_t *indeg, hipStream_t s);
/* tcol[pos[v]++] = u for every edge (u, v) of the split graph g. */
int ggak_transpose_fill(gga_dev_graph g, uint32_t *pos, uint32_t *tcol, hipStream_t s);
/* Workgroups of the cooperative bitmask BFS: per_cu per CU, within the occupancy bound. */
uint32_t ggak_msbfs_grid(int device, uint32_t per_cu);
/* Clears the bitmaps, seeds the k sources (d_src, device), builds level 0's push list. */
int ggak_msbfs_init(gga_dev_graph g, gga_dev_msbfs b, const uint32_t *d_src, hipStream_t s);
/* All levels in one cooperative launch (t: transposed graph with its chunk table). */
int ggak_msbfs_run(gga_dev_graph g, gga_dev_graph t, gga_dev_msbfs b, uint32_t max_levels, uint32_t grid,
                   uint32_t spin_limit_ms, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
