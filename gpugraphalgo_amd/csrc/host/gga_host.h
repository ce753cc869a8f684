/* gga_host.h -- internal declarations of the C host layer. */
#ifndef GGA_HOST_H
#define GGA_HOST_H
#include <stdint.h>
#include <stdio.h>

#include "../../../include/gga.h"

#include "../gga_kernels.h"

#define GGA_ERRBUF 512

/* What the engines need of a resident graph (graph.c); splits it with the
 * default threshold first if it has no chunk table. */
typedef struct gga_graph_view_t {
    gga_dev_graph dg;
    hipStream_t stream;
    int device;
    uint32_t split_gen;
    uint32_t maxdeg;
} gga_graph_view_t;
int gga_graph_view(gga_graph *g, gga_graph_view_t *v);
/* The graph's cached engine for gga_multi_sp (freed with the graph or a re-split). */
gga_ms **gga_graph_ms_slot(gga_graph *g);
/* Tuned defaults (tune.c) that traversals use when their opts leave them 0. */
void gga_graph_set_tuned(gga_graph *g, uint32_t grid, uint32_t dense_min, uint32_t layout);
uint32_t gga_graph_tuned_layout(const gga_graph *g);
/* mode < 0: no tuned mode; delta 0xFFFFFFFF: near-far tuned off, 0: untuned. */
void gga_graph_set_tuned_sssp(gga_graph *g, int32_t mode, uint32_t delta);
uint32_t gga_graph_wmean(const gga_graph *g);
/* Near-far width a query with opts.delta == 0 would use (0: off). */
uint32_t gga_graph_auto_delta(const gga_graph *g, int bfs);
int gga_graph_mesh_like(const gga_graph *g);   /* max degree within 4x the mean (+8) */
uint32_t gga_ms_step_grid(gga_ms *e);          /* an engine's default relax grid (0 on error) */
uint32_t gga_graph_engine_delta(const gga_graph *g);   /* engines' default near-far width (0: off) */
uint32_t gga_graph_maxdeg(const gga_graph *g);
/* Occupancy grid of the worklist kernel at `block` threads (0: the graph's size). */
uint32_t gga_graph_default_grid(gga_graph *g, int bfs, uint32_t block);
void gga_graph_set_tuned_block(gga_graph *g, uint32_t block);
/* State of gga_multi_bfs (msbfs.c), owned by the graph. */
void **gga_graph_msbfs_slot(gga_graph *g);
void gga_msbfs_release(void *state);

/* Cooperative launches (persistent kernels) of one process run one at a time:
 * each holds this lock from its launch until its stream has drained.  Two
 * such grids at once could each hold execution slots the other's grid
 * barrier waits for. */
void gga_coop_lock(void);
void gga_coop_unlock(void);

/* thread-local detail for gga_last_error() */
void gga_set_error(const char *fmt, ...);
void gga_clear_error(void);

/* 64-bit key radix sort (LSD, 16-bit digits over the low `bits` bits). */
int gga_radix_sort_u64(uint64_t *keys, uint64_t count, unsigned bits);
/* Builds CSR from sorted (src << 32 | dst) keys; keys must be sorted. */
int gga_csr_from_sorted_keys(uint32_t n, const uint64_t *keys, uint64_t count, gga_csr *out);

static inline uint64_t gga_splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

#endif
