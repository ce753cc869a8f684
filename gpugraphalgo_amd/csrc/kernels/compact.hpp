// compact.hpp -- list building after a flag step.
//
// A flag step records the next frontier as a per-vertex predicate (a step
// stamp, a lane mask); the chunk list is then rebuilt from it with no shared
// counter: count the chunk ids each tile of GGA_TILE vertices contributes,
// scan the tile counts, write.  Step launches and the cooperative kernels'
// steps run two passes, the write pass summing the counts before its own tile
// (compact_write_tiles); the scan as a pass of its own, by one workgroup
// (compact_scan), is left to the bitmask BFS's first list and the near-far
// refill.  Lists are in vertex order, so chunks of one vertex stay adjacent.
#pragma once
#include "../gga_dev.h"
#include "sync.hpp"
#include "wave.hpp"

namespace gga {

constexpr uint32_t CBLOCK = 256;                 // workgroup size of the compaction passes
constexpr uint32_t CITEMS = GGA_TILE / CBLOCK;   // vertices per thread per tile
constexpr uint32_t CSMALL = 4;                   // a thread writes at most this many ids itself

__device__ __forceinline__ void queue_store(uint32_t *qout, uint32_t pos, uint32_t id, uint32_t qcap,
                                            uint32_t *ctl) {
    // a vertex is listed at most once per step, so pos < C <= qcap; the guard
    // only turns a broken invariant into an error bit
    if (pos < qcap) qout[pos] = id;
    else __hip_atomic_fetch_or(&ctl[GGA_CTL_ERR], GGA_DEVERR_QFULL, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

// Append the wave's staged vertices to the chunk worklist: every staged vertex
// contributes its chunk range [cstart[v], cstart[v + 1]).
__device__ __forceinline__ void flush_pushes(const gga_dev_graph &g, const uint32_t *pbuf,
                                             uint32_t pcount, uint32_t *qout, uint32_t *tail,
                                             uint32_t qcap, uint32_t *ctl) {
    const uint32_t l = lane_id();
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i0 = 0; i0 < pcount; i0 += GGA_WAVE) {
        const uint32_t idx = i0 + l;
        uint32_t cs = 0, nch = 0;
        if (idx < pcount) {
            const uint32_t v = pbuf[idx];
            cs = g.cstart[v];
            nch = g.cstart[v + 1] - cs;
        }
        const uint32_t incl = wave_incl_scan(nch);
        const uint32_t tot = wave_last(incl);
        uint32_t base = 0;
        if (l == 0) base = atomic_add_u32(tail, tot);
        base = wave_bcast(base, 0);
        const uint32_t excl = incl - nch;
        for (uint32_t b = 0; b < tot; b += GGA_WAVE) {
            const uint32_t k = b + l;
            const uint32_t kk = k < tot ? k : tot - 1;
            const uint32_t j = wave_owner(incl, kk);
            const uint32_t id = wave_bcast(cs, j) + (kk - wave_bcast(excl, j));
            if (k < tot) queue_store(qout, base + k, id, qcap, ctl);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

struct ListOut {
    const uint32_t *cstart;  // first chunk of each vertex (n + 1)
    uint32_t *q;             // output list
    uint32_t qcap;           // its capacity
    uint32_t *tiles;         // per-tile counts, then offsets (GGA_TILES(n) entries)
    uint32_t *tail;          // list length
    uint32_t *ctl;           // error bits
};

// Per-thread view of one tile: CITEMS consecutive vertices v0 + tid * CITEMS + i.
// The list comes out in vertex order, so the next step's batches walk
// adjacent chunks (contiguous col / w runs across chunk boundaries), and a
// thread's ids go to consecutive positions.  Against vertices strided by
// CBLOCK: R-MAT s22 SSSP 2.24 -> 2.18 ms, the 4096² grid's persistent BFS /
// SSSP 9.5 / 15.4 -> 8.7 / 14.1 ms.
//
// A predicate with vec8 reads the marks of a thread's eight vertices with one
// or two vector loads (mark8: bit i for vertex v + i), and the marked
// vertices' chunk starts then come from two 16-byte loads of cstart and one
// word: a thread's vertices start at a multiple of 8, and the arrays are
// 256-byte aligned device allocations.  Threads whose vertices run past n
// take them one at a time.
static_assert(CITEMS == 8, "the vector path reads eight vertices per thread");
template <class Marked>
__device__ __forceinline__ uint32_t tile_items(const ListOut &o, uint32_t n, uint32_t t, const Marked &marked,
                                               uint32_t *cs, uint32_t *nch) {
    const uint64_t v0 = (uint64_t)t * GGA_TILE;
    const uint64_t vb = v0 + (uint64_t)threadIdx.x * CITEMS;
    uint32_t mine = 0;
    if constexpr (Marked::vec8) {
        if (vb + CITEMS <= n) {
            const uint32_t bits = marked.mark8(vb);
#pragma unroll
            for (uint32_t i = 0; i < CITEMS; i++) {
                cs[i] = 0;
                nch[i] = 0;
            }
            if (bits) {
                const uint4 a = *reinterpret_cast<const uint4 *>(o.cstart + vb);
                const uint4 b = *reinterpret_cast<const uint4 *>(o.cstart + vb + 4);
                const uint32_t c[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, o.cstart[vb + 8]};
#pragma unroll
                for (uint32_t i = 0; i < CITEMS; i++)
                    if ((bits >> i) & 1u) {
                        cs[i] = c[i];
                        nch[i] = c[i + 1] - c[i];
                    }
            }
#pragma unroll
            for (uint32_t i = 0; i < CITEMS; i++) mine += nch[i];
            return mine;
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < CITEMS; i++) {
        const uint64_t v = vb + i;
        cs[i] = 0;
        nch[i] = 0;
        if (v < n && marked((uint32_t)v)) {
            cs[i] = o.cstart[v];
            nch[i] = o.cstart[v + 1] - cs[i];
        }
        mine += nch[i];
    }
    return mine;
}

// pass 1, one tile: tiles[t] = number of chunk ids tile t contributes (also
// returned, to every thread)
template <class Marked>
__device__ __forceinline__ uint32_t compact_count_tile(const ListOut &o, uint32_t n, uint32_t t,
                                                       const Marked &marked, uint32_t *s_wave) {
    uint32_t cs[CITEMS], nch[CITEMS], excl;
    const uint32_t tot = block_excl_scan(tile_items(o, n, t, marked, cs, nch), excl, s_wave);
    if (threadIdx.x == 0) o.tiles[t] = tot;
    return tot;
}

// pass 2 (one workgroup): exclusive scan of tiles[]; the total is the new list
// length.  Each thread takes SCAN_ITEMS consecutive tiles (all loaded before
// the first is used), so a round covers CBLOCK * SCAN_ITEMS tiles with one
// workgroup scan: the 8 192 tiles of a 16.8 M-vertex graph take 4 rounds, not 32.
constexpr uint32_t SCAN_ITEMS = 8;
__device__ __forceinline__ void compact_scan(const ListOut &o, uint32_t n, uint32_t *s_wave) {
    const uint32_t ntiles = GGA_TILES(n);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += CBLOCK * SCAN_ITEMS) {
        const uint32_t i0 = base + threadIdx.x * SCAN_ITEMS;
        uint32_t x[SCAN_ITEMS], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
            x[k] = i0 + k < ntiles ? o.tiles[i0 + k] : 0u;
            sum += x[k];
        }
        uint32_t excl;
        const uint32_t tot = block_excl_scan(sum, excl, s_wave);
        uint32_t run = carry + excl;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
            if (i0 + k < ntiles) o.tiles[i0 + k] = run;
            run += x[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) *o.tail = carry;
}

// The chunk ids of one thread's vertices at list position pos.  A vertex with
// few chunks is written by its thread.  Longer runs (hubs' hundreds of chunk
// ids) are staged in LDS and written by every wave of the workgroup, 64 ids
// per round: in R-MAT the hubs sit at the lowest ids, so one tile -- and
// within it the first wave's vertices -- holds most of them, and written by
// their own wave they made that wave the whole pass's long pole (62 of the
// 154 us of an R-MAT s22 SSSP query's list writes, in one pass).  All threads of
// the workgroup must call this (it synchronises).
constexpr uint32_t CBIG = 256;   // long runs staged per round
__device__ __forceinline__ void write_items(const ListOut &o, const uint32_t *cs, const uint32_t *nch,
                                            uint32_t pos) {
    __shared__ uint32_t s_bpos[CBIG], s_bcs[CBIG], s_bcnt[CBIG];
    __shared__ uint32_t s_nb;
    uint32_t bpos[CITEMS];
    uint32_t pending = 0;   // bit i: item i is a long run not staged yet
#pragma unroll
    for (uint32_t i = 0; i < CITEMS; i++) {
        bpos[i] = pos;
        if (nch[i] <= CSMALL)
            for (uint32_t k = 0; k < nch[i]; k++) queue_store(o.q, pos + k, cs[i] + k, o.qcap, o.ctl);
        else
            pending |= 1u << i;
        pos += nch[i];
    }
    const uint32_t wv = threadIdx.x / GGA_WAVE, nw = blockDim.x / GGA_WAVE;
    for (;;) {
        if (threadIdx.x == 0) s_nb = 0;
        __syncthreads();
        while (pending) {
            const uint32_t i = (uint32_t)__builtin_ctz(pending);
            const uint32_t e = atomicAdd(&s_nb, 1u);
            if (e >= CBIG) break;   // full: stays pending for the next round
            uint32_t p = 0, c = 0, k = 0;
#pragma unroll
            for (uint32_t j = 0; j < CITEMS; j++)
                if (j == i) {
                    p = bpos[j];
                    c = cs[j];
                    k = nch[j];
                }
            s_bpos[e] = p;
            s_bcs[e] = c;
            s_bcnt[e] = k;
            pending &= pending - 1u;
        }
        __syncthreads();
        const uint32_t nb = s_nb < CBIG ? s_nb : CBIG;
        for (uint32_t e = wv; e < nb; e += nw) {
            const uint32_t p = s_bpos[e], c0 = s_bcs[e], cnt = s_bcnt[e];
            for (uint32_t k = lane_id(); k < cnt; k += GGA_WAVE) queue_store(o.q, p + k, c0 + k, o.qcap, o.ctl);
        }
        if (!__syncthreads_or(pending != 0u)) break;
    }
}

// pass 3, one tile: write
template <class Marked>
__device__ __forceinline__ void compact_write_tile(const ListOut &o, uint32_t n, uint32_t t, const Marked &marked,
                                                   uint32_t *s_wave) {
    uint32_t cs[CITEMS], nch[CITEMS], excl;
    const uint32_t mine = tile_items(o, n, t, marked, cs, nch);
    block_excl_scan(mine, excl, s_wave);
    write_items(o, cs, nch, o.tiles[t] + excl);
}

// Passes 2 and 3 in one launch (every workgroup of the grid calls this):
// tiles[] still holds the counts, and each workgroup sums the counts of the
// tiles before its own -- at most GGA_TILES(n) L2-resident words per
// workgroup, cheaper than the single-workgroup scan pass's launch, which
// costs a kernel boundary (~4.5 us on gfx950) per list.  The workgroup of
// the last tile stores the list length.
template <class Marked>
__device__ __forceinline__ void compact_write_tiles(const ListOut &o, uint32_t n, const Marked &marked,
                                                    uint32_t *s_wave) {
    const uint32_t ntiles = (uint32_t)GGA_TILES((uint64_t)n);
    uint32_t off = 0, upto = 0;   // off = tiles[0] + ... + tiles[upto - 1]
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t part = 0, ex;
        for (uint32_t i = upto + threadIdx.x; i < t; i += blockDim.x) part += o.tiles[i];
        off += block_excl_scan(part, ex, s_wave);
        upto = t;
        uint32_t cs[CITEMS], nch[CITEMS], excl;
        const uint32_t mine = tile_items(o, n, t, marked, cs, nch);
        const uint32_t tot = block_excl_scan(mine, excl, s_wave);
        write_items(o, cs, nch, off + excl);
        if (t == ntiles - 1u && threadIdx.x == 0) *o.tail = off + tot;
    }
}

}  // namespace gga
