// compact.hpp -- list building after a flag step.
//
// A flag step records the next frontier as a per-vertex predicate (a step
// stamp, a lane mask); three passes then rebuild the chunk list from it with
// no shared counter: count the chunk ids each tile of GGA_TILE vertices
// contributes, scan the tile counts (one workgroup), write.  Lists are in
// vertex order, so chunks of one vertex stay adjacent.
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
        const uint32_t tot = wave_bcast(incl, GGA_WAVE - 1);
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

// Per-thread view of one tile: CITEMS vertices v0 + i * CBLOCK + tid (coalesced).
template <class Marked>
__device__ __forceinline__ uint32_t tile_items(const ListOut &o, uint32_t n, uint32_t t, const Marked &marked,
                                               uint32_t *cs, uint32_t *nch) {
    const uint64_t v0 = (uint64_t)t * GGA_TILE;
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t i = 0; i < CITEMS; i++) {
        const uint64_t v = v0 + (uint64_t)i * CBLOCK + threadIdx.x;
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

// pass 1, one tile: tiles[t] = number of chunk ids tile t contributes
template <class Marked>
__device__ __forceinline__ void compact_count_tile(const ListOut &o, uint32_t n, uint32_t t, const Marked &marked,
                                                   uint32_t *s_wave) {
    uint32_t cs[CITEMS], nch[CITEMS], excl;
    const uint32_t tot = block_excl_scan(tile_items(o, n, t, marked, cs, nch), excl, s_wave);
    if (threadIdx.x == 0) o.tiles[t] = tot;
}

// pass 2 (one workgroup): exclusive scan of tiles[]; the total is the new list length
__device__ __forceinline__ void compact_scan(const ListOut &o, uint32_t n, uint32_t *s_wave) {
    const uint32_t ntiles = GGA_TILES(n);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += CBLOCK) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < ntiles ? o.tiles[i] : 0u;
        uint32_t excl;
        const uint32_t tot = block_excl_scan(x, excl, s_wave);
        if (i < ntiles) o.tiles[i] = carry + excl;
        carry += tot;
    }
    if (threadIdx.x == 0) *o.tail = carry;
}

// The chunk ids of one thread's vertices at list position pos.  A vertex with
// few chunks is written by its thread; a hub's hundreds of chunk ids are
// written by the whole wave, 64 per round.
__device__ __forceinline__ void write_items(const ListOut &o, const uint32_t *cs, const uint32_t *nch,
                                            uint32_t pos) {
    uint32_t bpos[CITEMS];
#pragma unroll
    for (uint32_t i = 0; i < CITEMS; i++) {
        bpos[i] = pos;
        if (nch[i] <= CSMALL)
            for (uint32_t k = 0; k < nch[i]; k++) queue_store(o.q, pos + k, cs[i] + k, o.qcap, o.ctl);
        pos += nch[i];
    }
#pragma unroll
    for (uint32_t i = 0; i < CITEMS; i++) {
        uint64_t big = __ballot(nch[i] > CSMALL);
        while (big) {
            const uint32_t j = (uint32_t)__builtin_ctzll(big);
            big &= big - 1;
            const uint32_t p = wave_bcast(bpos[i], j), c0 = wave_bcast(cs[i], j);
            const uint32_t cnt = wave_bcast(nch[i], j);
            for (uint32_t k = lane_id(); k < cnt; k += GGA_WAVE) queue_store(o.q, p + k, c0 + k, o.qcap, o.ctl);
        }
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

// ---- one pass: count, offset and write in a single launch (decoupled
// look-back, Merrill & Garland 2016).  Workgroup t owns tile t: it counts,
// publishes its count, sums the counts of the tiles before it by looking back
// over their published states, publishes its inclusive total and writes.  A
// tile waits only on tiles of lower index, which the dispatcher has started
// before it, so every wait ends; the wait is bounded by the clock anyway.
// Tile state words: bits 63-62 flag (0 none, 1 count, 2 inclusive total),
// 61-32 the compaction's epoch, 31-0 the value.  The epoch (a ctl word) goes
// up by one per compaction, so states of the previous one read as "none" and
// the array never needs clearing.
constexpr uint64_t TS_AGG = 1ull << 62;
constexpr uint64_t TS_INC = 2ull << 62;
constexpr uint64_t LOOKBACK_SPIN_TICKS = 500000000ull;   // 5 s of the 100 MHz wall clock

__device__ __forceinline__ uint64_t ts_pack(uint32_t epoch, uint64_t flag, uint32_t v) {
    return flag | ((uint64_t)(epoch & 0x3FFFFFFFu) << 32) | v;
}

// Wave 0: the sum of the values of tiles [0, t).
__device__ __forceinline__ uint32_t lookback(const uint64_t *ts, uint32_t t, uint32_t epoch, uint32_t *ctl) {
    const uint32_t l = lane_id();
    const uint64_t t0 = wall_clock64();
    uint32_t prefix = 0;
    int64_t j = (int64_t)t - 1;
    while (j >= 0) {
        // lane l reads tile j - l; tiles before 0 read as an inclusive total of 0
        const int64_t idx = j - (int64_t)l;
        const uint64_t wd = idx >= 0 ? __hip_atomic_load(&ts[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : ts_pack(epoch, TS_INC, 0u);
        const bool ready = (wd >> 62) != 0u && (uint32_t)((wd >> 32) & 0x3FFFFFFFu) == (epoch & 0x3FFFFFFFu);
        const uint64_t nr = __ballot(!ready), inc = __ballot(ready && (wd >> 62) == 2u);
        const uint32_t k_inc = inc ? (uint32_t)__builtin_ctzll(inc) : GGA_WAVE;
        const uint32_t k_nr = nr ? (uint32_t)__builtin_ctzll(nr) : GGA_WAVE;
        if (k_nr < k_inc) {   // a tile between us and the nearest total has not counted yet
            if (__ballot(l == 0 && wall_clock64() - t0 > LOOKBACK_SPIN_TICKS)) {
                if (l == 0) __hip_atomic_fetch_or(&ctl[GGA_CTL_ERR], GGA_DEVERR_SPIN, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        prefix += wave_sum(l <= k_inc ? (uint32_t)wd : 0u);
        if (k_inc < GGA_WAVE) break;
        j -= GGA_WAVE;
    }
    return prefix;
}

// Launch with exactly GGA_TILES(n) workgroups of CBLOCK threads, all taking
// part (the caller returns from every workgroup or from none).
template <class Marked>
__device__ __forceinline__ void compact_onepass(const ListOut &o, uint32_t n, const Marked &marked, uint64_t *ts,
                                                uint32_t *epoch_word, uint32_t *s_wave, uint32_t *s_pre) {
    const uint32_t t = blockIdx.x, ntiles = GGA_TILES(n);
    const uint32_t epoch = load_relaxed(epoch_word);
    uint32_t cs[CITEMS], nch[CITEMS], excl;
    const uint32_t tot = block_excl_scan(tile_items(o, n, t, marked, cs, nch), excl, s_wave);
    if (threadIdx.x == 0)
        __hip_atomic_store(&ts[t], ts_pack(epoch, t == 0 ? TS_INC : TS_AGG, tot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < GGA_WAVE) {
        const uint32_t p = t ? lookback(ts, t, epoch, o.ctl) : 0u;
        if (threadIdx.x == 0) {
            s_pre[0] = p;
            if (t) __hip_atomic_store(&ts[t], ts_pack(epoch, TS_INC, p + tot), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            // the last tile has seen every tile publish, so every workgroup
            // has read the epoch: the next compaction may use the next one
            if (t == ntiles - 1u) {
                *o.tail = p + tot;
                __hip_atomic_store(epoch_word, epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    write_items(o, cs, nch, s_pre[0] + excl);
}

}  // namespace gga
