// msbfs.hip -- multi-source BFS with bitmask lanes: up to 256 BFS per launch.
//
// No reference counterpart (the reference runs one BFS per process,
// bfs_sssp.cu:203); each lane's levels equal the single-source BFS of the
// reference (graph.h:150-184 / M_my_drelax with unit weights).  Level-
// synchronous: at level L every lane that reached a vertex at level L passes
// it on along every out-edge, so a lane's bit first appears at v at the
// level equal to its BFS distance.
//
//   push level (small frontier): the chunks of frontier vertices (one list,
//     built by the shared compaction) OR their front bits into next[v] of each
//     destination, masked by the destination's seen bits; 4 lanes of a wave
//     per edge, one per 64-bit word.
//   pull level (large frontier): every in-chunk of the transposed graph whose
//     owner still misses lanes ORs the front bits of its in-neighbours until
//     the missing lanes are all found (direction-optimising BFS, Beamer et al.
//     2012, on bitmaps as in multi-source BFS, Then et al. 2014).
//   finalize: new = next & ~seen; seen |= new; front = new.
// The fronts of the first nsnap levels stay in slots of their own (front
// slot s holds the lanes discovered at level s) and one pass at the end
// writes every level row from them and from seen, each output line once:
// level rows are lane-major, and writing each level's new bits as they
// appear dirtied every line of them again at every wide level (256 BFS on
// R-MAT s20: about 1.1 ms of 3.3 in the writes).  Levels from nsnap on
// share one slot and write their values at discovery.
// One cooperative launch runs all levels; grid barriers separate the phases,
// and every workgroup sums the level's partials and so takes the same
// push / pull / end decision without a further barrier.
#include "../gga_kernels.h"
#include "compact.hpp"
#include "sync.hpp"
#include "wave.hpp"

using namespace gga;

namespace {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t WAVES = BLOCK / GGA_WAVE;
constexpr uint32_t WORDS = 4;            // 64-bit words per vertex bitmap
constexpr uint32_t GROUPS = GGA_WAVE / WORDS;  // edges (or chunks) per wave round
#ifndef GGA_MSBFS_UNROLL
#define GGA_MSBFS_UNROLL 4
#endif
constexpr uint32_t UNROLL = GGA_MSBFS_UNROLL;

__device__ __forceinline__ void atomic_or_u64(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// bits of the lanes in use within word j
__device__ __forceinline__ uint64_t lane_mask(uint32_t k, uint32_t j) {
    const uint32_t lo = j * 64u;
    if (k >= lo + 64u) return ~0ull;
    if (k <= lo) return 0ull;
    return (1ull << (k - lo)) - 1ull;
}

struct FrontMarked {
    const uint64_t *front;
    static constexpr bool vec8 = false;
    __device__ __forceinline__ bool operator()(uint32_t v) const {
        const uint64_t *f = front + (uint64_t)v * WORDS;
        return (f[0] | f[1] | f[2] | f[3]) != 0ull;
    }
};

__device__ __forceinline__ ListOut bfs_list(const gga_dev_graph &g, const gga_dev_msbfs &b) {
    ListOut o;
    o.cstart = g.cstart;
    o.q = b.q;
    o.qcap = b.qcap;
    o.tiles = b.tiles;
    o.tail = &b.ctl[GGA_CTL_TAIL(0)];
    o.ctl = b.ctl;
    return o;
}

// ------------------------------------------------------------- push level
// front slot of level L: its own below nsnap, the shared one from there on
__device__ __forceinline__ uint64_t *front_at(const gga_dev_msbfs &b, uint32_t n, uint32_t level) {
    return b.front + (uint64_t)(level < b.nsnap ? level : b.nsnap) * n * WORDS;
}

__device__ __forceinline__ void push_level(const gga_dev_graph &g, const gga_dev_msbfs &b, const uint64_t *front,
                                           uint32_t nq) {
    const uint32_t waves = gridDim.x * WAVES;
    const uint32_t wid = blockIdx.x * WAVES + threadIdx.x / GGA_WAVE;
    uint32_t bsz = nq / (8u * waves);   // ~8 batches per wave (16 measured no faster here)
    bsz = bsz < 1u ? 1u : (bsz > GGA_WAVE ? GGA_WAVE : bsz);
    const uint32_t nbatch = (nq + bsz - 1) / bsz;
    const uint32_t l = lane_id(), grp = l / WORDS, j = l % WORDS;
    for (uint32_t bt = wid; bt < nbatch; bt += waves) {
        const uint32_t idx = bt * bsz + l;
        const bool valid = l < bsz && idx < nq;
        uint32_t eb = 0, ne = 0, u = 0;
        if (valid) {
            const uint32_t c = stream_load(&b.q[idx]);
            u = stream_load(&g.cown[c]);
            eb = stream_load(&g.cbeg[c]);
            ne = stream_load(&g.cbeg[c + 1]) - eb;
        }
        const uint32_t incl = wave_incl_scan(ne);
        const uint32_t total = wave_last(incl);
        const uint32_t excl = incl - ne;
        for (uint32_t base = 0; base < total; base += GROUPS * UNROLL) {
            uint32_t v[UNROLL];
            uint64_t f[UNROLL], sn[UNROLL];
            bool act[UNROLL];
#pragma unroll
            for (uint32_t i = 0; i < UNROLL; i++) {
                const uint32_t k = base + i * GROUPS + grp;
                act[i] = k < total;
                const uint32_t kk = act[i] ? k : total - 1u;
                const uint32_t o = wave_owner(incl, kk);
                const uint32_t e = wave_bcast(eb, o) + (kk - wave_bcast(excl, o));
                const uint32_t uu = wave_bcast(u, o);
                v[i] = stream_load(&g.col[e]);
                f[i] = front[(uint64_t)uu * WORDS + j];
            }
#pragma unroll
            for (uint32_t i = 0; i < UNROLL; i++) sn[i] = b.seen[(uint64_t)v[i] * WORDS + j];
#pragma unroll
            for (uint32_t i = 0; i < UNROLL; i++) {
                const uint64_t x = f[i] & ~sn[i];
                if (act[i] && x) atomic_or_u64(&b.next[(uint64_t)v[i] * WORDS + j], x);
            }
        }
    }
}

// ------------------------------------------------------------- pull level
// t: transposed graph with its own chunk table (in-edges of each vertex).
__device__ __forceinline__ void pull_level(const gga_dev_graph &t, const gga_dev_msbfs &b, const uint64_t *front) {
    const uint32_t l = lane_id(), grp = l / WORDS, j = l % WORDS;
    const uint64_t lm = lane_mask(b.k, j);
    const uint64_t gstride = (uint64_t)gridDim.x * WAVES * GROUPS;
    for (uint64_t c = ((uint64_t)blockIdx.x * WAVES + threadIdx.x / GGA_WAVE) * GROUPS + grp; c < t.C;
         c += gstride) {
        const uint32_t v = t.cown[c];
        const uint64_t miss = ~b.seen[(uint64_t)v * WORDS + j] & lm;
        if (!miss) continue;
        const uint32_t eb = t.cbeg[c], ee = t.cbeg[c + 1];
        uint64_t acc = 0;
        uint32_t e = eb;
        for (; e + UNROLL <= ee && (acc & miss) != miss; e += UNROLL) {
            uint32_t u[UNROLL];
#pragma unroll
            for (uint32_t i = 0; i < UNROLL; i++) u[i] = t.col[e + i];
#pragma unroll
            for (uint32_t i = 0; i < UNROLL; i++) acc |= front[(uint64_t)u[i] * WORDS + j];
        }
        for (; e < ee && (acc & miss) != miss; e++) acc |= front[(uint64_t)t.col[e] * WORDS + j];
        acc &= miss;
        if (acc) atomic_or_u64(&b.next[(uint64_t)v * WORDS + j], acc);
    }
}

// Level values of the set bits of `bits` (word j of vertex v), uniform loop.
// The wave walks only the bit positions set in some lane's word (their OR,
// a butterfly over the wave, kept scalar), not all 64: one coalesced store
// per position for the lanes that hold it.
__device__ __forceinline__ void write_levels(const gga_dev_msbfs &b, uint32_t n, uint32_t v, bool valid,
                                             uint32_t j, uint64_t bits, uint32_t value) {
    const uint64_t mine = valid ? bits : 0ull;
    uint32_t lo = (uint32_t)mine, hi = (uint32_t)(mine >> 32);
#pragma unroll
    for (uint32_t o = 32; o >= 1; o >>= 1) {
        lo |= __shfl_xor(lo, (int)o, GGA_WAVE);
        hi |= __shfl_xor(hi, (int)o, GGA_WAVE);
    }
    uint64_t any = (uint64_t)__builtin_amdgcn_readfirstlane(hi) << 32 | __builtin_amdgcn_readfirstlane(lo);
    while (any) {
        const uint32_t t = (uint32_t)__builtin_ctzll(any);
        any &= any - 1u;
        const uint32_t lane = j * 64u + t;
        if (lane >= b.k) break;
        if ((mine >> t) & 1ull) b.lev[(uint64_t)lane * n + v] = value;
    }
}

// the four words of vertex v's bitmap, as two 16-byte accesses (bitmaps are
// 32-byte aligned per vertex)
__device__ __forceinline__ void load_bits(const uint64_t *p, uint32_t v, uint64_t w[WORDS]) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p + (uint64_t)v * WORDS);
    const uint4 a = q[0], c = q[1];
    w[0] = (uint64_t)a.x | (uint64_t)a.y << 32;
    w[1] = (uint64_t)a.z | (uint64_t)a.w << 32;
    w[2] = (uint64_t)c.x | (uint64_t)c.y << 32;
    w[3] = (uint64_t)c.z | (uint64_t)c.w << 32;
}
__device__ __forceinline__ void store_bits(uint64_t *p, uint32_t v, const uint64_t w[WORDS]) {
    uint4 *q = reinterpret_cast<uint4 *>(p + (uint64_t)v * WORDS);
    q[0] = make_uint4((uint32_t)w[0], (uint32_t)(w[0] >> 32), (uint32_t)w[1], (uint32_t)(w[1] >> 32));
    q[1] = make_uint4((uint32_t)w[2], (uint32_t)(w[2] >> 32), (uint32_t)w[3], (uint32_t)(w[3] >> 32));
}

struct LevelSums {
    uint64_t nf = 0, mf = 0, reached = 0, mreach = 0;
};

// finalize one tile into the front slot of level + 1; returns the tile's
// push-list chunk count in tiles[t]
__device__ __forceinline__ void finalize_tile(const gga_dev_graph &g, const gga_dev_msbfs &b, uint64_t *nfront,
                                              uint32_t t, uint32_t level, LevelSums &ls, uint32_t *s_wave) {
    const bool direct = level + 1u >= b.nsnap;
    const uint64_t v0 = (uint64_t)t * GGA_TILE;
    uint32_t chunks = 0;
    for (uint32_t i = 0; i < CITEMS; i++) {
        const uint64_t vv = v0 + (uint64_t)i * CBLOCK + threadIdx.x;
        const bool valid = vv < g.n;
        const uint32_t v = valid ? (uint32_t)vv : 0u;
        // a vertex's four words move as two 16-byte accesses
        uint64_t nx[WORDS] = {0, 0, 0, 0}, sn[WORDS] = {0, 0, 0, 0}, nw[WORDS];
        if (valid) {
            load_bits(b.next, v, nx);
            load_bits(b.seen, v, sn);
        }
        uint64_t any = 0, anx = 0;
        uint32_t pc = 0;
#pragma unroll
        for (uint32_t j = 0; j < WORDS; j++) {
            nw[j] = nx[j] & ~sn[j];
            any |= nw[j];
            anx |= nx[j];
            pc += (uint32_t)__popcll(nw[j]);
        }
        if (valid) {
            if (anx) {
                const uint64_t z[WORDS] = {0, 0, 0, 0};
                store_bits(b.next, v, z);
            }
            if (any) {
                uint64_t ns[WORDS];
#pragma unroll
                for (uint32_t j = 0; j < WORDS; j++) ns[j] = sn[j] | nw[j];
                store_bits(b.seen, v, ns);
            }
            store_bits(nfront, v, nw);
        }
        if (direct)
#pragma unroll
            for (uint32_t j = 0; j < WORDS; j++) write_levels(b, g.n, v, valid, j, nw[j], level + 1u);
        if (valid && any) {
            const uint32_t deg = g.row[v + 1] - g.row[v];
            ls.nf += 1;
            ls.mf += deg;
            ls.reached += pc;
            ls.mreach += (uint64_t)pc * deg;
            chunks += g.cstart[v + 1] - g.cstart[v];
        }
    }
    uint32_t excl;
    const uint32_t tot = block_excl_scan(chunks, excl, s_wave);
    if (threadIdx.x == 0) b.tiles[t] = tot;
}

// The level rows, each (vertex, lane) written once: the level of its front
// slot (slots 0 .. nslots-1; a lane's bit is in at most one, and the slot
// index is gathered bit-sliced: s0 holds the bits of odd slots, s1 of slots
// with bit 1 set, ...), inf where the lane never reached the vertex, nothing
// where it did at a level from nsnap on (written at discovery).  A vertex
// no lane reached reads no slot.
__device__ __forceinline__ void write_rows(const gga_dev_graph &g, const gga_dev_msbfs &b, uint32_t nslots) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t nn = ((uint64_t)g.n + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole waves loop
    const uint64_t slot_words = (uint64_t)g.n * WORDS;
    for (uint64_t vv = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; vv < nn; vv += stride) {
        const bool valid = vv < g.n;
        const uint32_t v = valid ? (uint32_t)vv : 0u;
        uint64_t sn[WORDS] = {0, 0, 0, 0}, found[WORDS] = {0, 0, 0, 0}, s0[WORDS] = {0, 0, 0, 0},
                 s1[WORDS] = {0, 0, 0, 0}, s2[WORDS] = {0, 0, 0, 0}, s3[WORDS] = {0, 0, 0, 0};
        if (valid) {
            load_bits(b.seen, v, sn);
            if (sn[0] | sn[1] | sn[2] | sn[3])
                for (uint32_t s = 0; s < nslots; s++) {
                    uint64_t w[WORDS];
                    load_bits(b.front + s * slot_words, v, w);
#pragma unroll
                    for (uint32_t j = 0; j < WORDS; j++) {
                        found[j] |= w[j];
                        s0[j] |= (s & 1u) ? w[j] : 0ull;
                        s1[j] |= (s & 2u) ? w[j] : 0ull;
                        s2[j] |= (s & 4u) ? w[j] : 0ull;
                        s3[j] |= (s & 8u) ? w[j] : 0ull;
                    }
                }
        }
#pragma unroll
        for (uint32_t j = 0; j < WORDS; j++) {
            if (j * 64u >= b.k) break;
            const uint64_t put = found[j] | (valid ? ~sn[j] & lane_mask(b.k, j) : 0ull);
            // rows are not read back here: non-temporal stores (2.08 -> 2.06 ms
            // for 256 BFS on R-MAT s20)
            for (uint32_t t = 0; t < 64u && j * 64u + t < b.k; t++) {
                if (!((put >> t) & 1ull)) continue;
                const uint32_t lv = (uint32_t)((s0[j] >> t) & 1ull) | (uint32_t)((s1[j] >> t) & 1ull) << 1 |
                                    (uint32_t)((s2[j] >> t) & 1ull) << 2 | (uint32_t)((s3[j] >> t) & 1ull) << 3;
                __builtin_nontemporal_store(((found[j] >> t) & 1ull) ? lv : b.inf,
                                            &b.lev[(uint64_t)(j * 64u + t) * g.n + v]);
            }
        }
    }
}

__device__ __forceinline__ uint64_t block_sum(uint64_t x, uint64_t *s_red) {
    x = wave_sum64(x);
    if (lane_id() == 0) s_red[threadIdx.x / GGA_WAVE] = x;
    __syncthreads();
    uint64_t s = 0;
    for (uint32_t i = 0; i < WAVES; i++) s += s_red[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------- kernels
__global__ __launch_bounds__(BLOCK) void k_msbfs_clear(gga_dev_graph g, gga_dev_msbfs b) {
    const uint64_t total = (uint64_t)g.n * WORDS;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        b.seen[i] = 0ull;
        b.front[i] = 0ull;
        b.next[i] = 0ull;
    }
    // every control word but the barrier generation (each launch reads that
    // at its start): error bits and the arrival counts of both barrier levels
    // start at zero, also after a launch that left a barrier on a timeout
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < GGA_CTL_WORDS; i += blockDim.x)
            if (i != GGA_CTL_BARGEN) b.ctl[i] = 0u;
}

// lane i starts at src[i]: level 0 there; reached / out-edge totals seeded
__global__ __launch_bounds__(BLOCK) void k_msbfs_seed(gga_dev_graph g, gga_dev_msbfs b, const uint32_t *src) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.k) return;
    const uint32_t x = src[i];
    const uint64_t bit = 1ull << (i % 64u);
    atomic_or_u64(&b.seen[(uint64_t)x * WORDS + i / 64u], bit);
    atomic_or_u64(&b.front[(uint64_t)x * WORDS + i / 64u], bit);
    if (b.nsnap == 0) b.lev[(uint64_t)i * g.n + x] = 0u;  // else from front slot 0 at the end
    uint64_t *tot = reinterpret_cast<uint64_t *>(b.ctl + GGA_CTL_CNT);
    atomic_add_u64(&tot[0], 1ull);
    atomic_add_u64(&tot[1], (uint64_t)(g.row[x + 1] - g.row[x]));
}

__global__ __launch_bounds__(CBLOCK) void k_msbfs_list_count(gga_dev_graph g, gga_dev_msbfs b) {
    __shared__ uint32_t s_wave[CBLOCK / GGA_WAVE];
    const ListOut o = bfs_list(g, b);
    for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x)
        compact_count_tile(o, g.n, t, FrontMarked{b.front}, s_wave);
}
__global__ __launch_bounds__(CBLOCK) void k_msbfs_list_scan(gga_dev_graph g, gga_dev_msbfs b) {
    __shared__ uint32_t s_wave[CBLOCK / GGA_WAVE];
    compact_scan(bfs_list(g, b), g.n, s_wave);
}
__global__ __launch_bounds__(CBLOCK) void k_msbfs_list_write(gga_dev_graph g, gga_dev_msbfs b) {
    __shared__ uint32_t s_wave[CBLOCK / GGA_WAVE];
    const ListOut o = bfs_list(g, b);
    for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x)
        compact_write_tile(o, g.n, t, FrontMarked{b.front}, s_wave);
}

__global__ __launch_bounds__(BLOCK) void k_msbfs(gga_dev_graph g, gga_dev_graph t, gga_dev_msbfs b,
                                                 uint32_t max_levels, uint64_t spin_ticks) {
    __shared__ uint32_t s_wave[WAVES];
    __shared__ uint64_t s_red[WAVES];
    __shared__ uint32_t s_state[3];
    uint32_t gen = 0;
    if (threadIdx.x == 0)
        gen = __hip_atomic_load(&b.ctl[GGA_CTL_BARGEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const ListOut o = bfs_list(g, b);
    const uint32_t ntiles = GGA_TILES(g.n);
    uint32_t level = 0, filled = 1;   // front slots written: 0 (the sources), then one per level
    uint32_t dir = 0;                 // the first level pushes from the sources' list
    for (; level < max_levels; level++) {
        if (threadIdx.x == 0) {
            s_state[1] = b.ctl[GGA_CTL_TAIL(0)];
            s_state[2] = __hip_atomic_load(&b.ctl[GGA_CTL_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t nq = s_state[1] < b.qcap ? s_state[1] : b.qcap, err = s_state[2];
        __syncthreads();
        if (err) break;
        const uint64_t *front = front_at(b, g.n, level);
        uint64_t *nfront = front_at(b, g.n, level + 1u);
        if (dir) pull_level(t, b, front);
        else push_level(g, b, front, nq);
        grid_barrier(b.ctl, gridDim.x, gen, spin_ticks);
        LevelSums ls;
        for (uint32_t tt = blockIdx.x; tt < ntiles; tt += gridDim.x)
            finalize_tile(g, b, nfront, tt, level, ls, s_wave);
        filled = level + 2u;
        const uint64_t nf = block_sum(ls.nf, s_red), mf = block_sum(ls.mf, s_red);
        const uint64_t rs = block_sum(ls.reached, s_red), mr = block_sum(ls.mreach, s_red);
        if (threadIdx.x == 0) {
            b.part[4u * blockIdx.x + 0] = nf;
            b.part[4u * blockIdx.x + 1] = mf;
            b.part[4u * blockIdx.x + 2] = rs;
            b.part[4u * blockIdx.x + 3] = mr;
        }
        grid_barrier(b.ctl, gridDim.x, gen, spin_ticks);
        // every workgroup sums the partials (exact integer sums: the same
        // decision everywhere) rather than waiting on one workgroup's
        // decision behind another grid barrier; the partials are rewritten
        // only after the next level's first barrier, which every workgroup
        // reaches after reading them
        uint64_t x[4] = {0, 0, 0, 0};
        for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x)
            for (uint32_t k = 0; k < 4; k++) x[k] += b.part[4u * i + k];
        for (uint32_t k = 0; k < 4; k++) x[k] = block_sum(x[k], s_red);
        const bool pull = b.alpha && x[1] * (uint64_t)b.alpha > (uint64_t)g.m;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            uint64_t *tot = reinterpret_cast<uint64_t *>(b.ctl + GGA_CTL_CNT);
            tot[0] += x[2];
            tot[1] += x[3];
            b.ctl[GGA_CTL_DONE] = x[0] == 0 ? 1u : 0u;
            b.ctl[GGA_CTL_DIR] = pull ? 1u : 0u;
            b.ctl[GGA_CTL_STEPS] = level + 1u;
            if (pull) b.ctl[GGA_CTL_NPULL] += 1u;
        }
        if (x[0] == 0) break;
        if (!pull) {
            // the list of the next level: each workgroup sums the counts of
            // the tiles before its own, the last tile's stores the length
            compact_write_tiles(o, g.n, FrontMarked{nfront}, s_wave);
            grid_barrier(b.ctl, gridDim.x, gen, spin_ticks);
        }
        dir = pull ? 1u : 0u;
    }
    write_rows(g, b, filled < b.nsnap ? filled : b.nsnap);
}

// ------------------------------------------------------ transposed graph
__global__ __launch_bounds__(BLOCK) void k_indeg(const uint32_t *col, uint32_t m, uint32_t *indeg) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (uint64_t)gridDim.x * blockDim.x)
        atomic_add_u32(&indeg[col[e]], 1u);
}

// one thread per forward chunk whose owner's out-degree is in [lo, hi) (hi 0:
// no bound); the
// in-edge order within one such pass is arbitrary
__global__ __launch_bounds__(BLOCK) void k_transpose_fill(gga_dev_graph g, uint32_t *pos, uint32_t *tcol, uint32_t lo,
                                                          uint32_t hi) {
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g.C;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = g.cown[c];
        const uint32_t d = g.row[u + 1] - g.row[u];
        if (d < lo || (hi != 0u && d >= hi)) continue;
        for (uint32_t e = g.cbeg[c]; e < g.cbeg[c + 1]; e++) tcol[atomic_add_u32(&pos[g.col[e]], 1u)] = u;
    }
}

inline int launch_status() { return (int)hipGetLastError(); }

inline uint32_t grid_for(uint64_t items, uint32_t cap) {
    uint64_t b = (items + BLOCK - 1) / BLOCK;
    if (b == 0) b = 1;
    return (uint32_t)(b < cap ? b : cap);
}

inline uint32_t compact_grid(uint32_t n) {
    const uint32_t ntiles = GGA_TILES(n);
    return ntiles < 2048u ? (ntiles ? ntiles : 1u) : 2048u;
}

}  // namespace

extern "C" int ggak_indeg(const uint32_t *col, uint32_t m, uint32_t *indeg, hipStream_t s) {
    if (m) k_indeg<<<grid_for(m, 8192), BLOCK, 0, s>>>(col, m, indeg);
    return launch_status();
}

// In-edges from sources of higher out-degree first: one pass per power-of-two
// degree class, from the top.  Hubs are reached by most lanes within the
// first levels, so a pull level finds a vertex's missing lanes after fewer
// gathers.
extern "C" int ggak_transpose_fill(gga_dev_graph g, uint32_t *pos, uint32_t *tcol, hipStream_t s) {
    for (int cls = 31; cls >= 0; cls--) {
        const uint32_t lo = 1u << cls, hi = cls == 31 ? 0u : 2u << cls;
        if (lo > g.m && cls > 0) continue;
        k_transpose_fill<<<grid_for(g.C, 8192), BLOCK, 0, s>>>(g, pos, tcol, lo, hi);
    }
    return launch_status();
}

extern "C" uint32_t ggak_msbfs_grid(int device, uint32_t per_cu) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void *>(&k_msbfs), BLOCK, 0) !=
            hipSuccess ||
        per <= 0)
        return 0;
    // the cooperative launch keeps well under the co-residency the occupancy
    // query reports (6 per CU for this kernel)
    if (per_cu == 0) per_cu = 1;
    const uint32_t pc = (uint32_t)per < per_cu ? (uint32_t)per : per_cu;
    return (uint32_t)cus * pc;
}

extern "C" int ggak_msbfs_init(gga_dev_graph g, gga_dev_msbfs b, const uint32_t *d_src, hipStream_t s) {
    if (b.k == 0 || b.k > GGA_MSBFS_MAX || b.nsnap > GGA_MSBFS_SNAP_MAX) return (int)hipErrorInvalidValue;
    k_msbfs_clear<<<grid_for((uint64_t)g.n * WORDS, 8192), BLOCK, 0, s>>>(g, b);
    k_msbfs_seed<<<(b.k + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(g, b, d_src);
    const uint32_t cg = compact_grid(g.n);
    k_msbfs_list_count<<<cg, CBLOCK, 0, s>>>(g, b);
    k_msbfs_list_scan<<<1, CBLOCK, 0, s>>>(g, b);
    k_msbfs_list_write<<<cg, CBLOCK, 0, s>>>(g, b);
    return launch_status();
}

extern "C" int ggak_msbfs_run(gga_dev_graph g, gga_dev_graph t, gga_dev_msbfs b, uint32_t max_levels,
                              uint32_t grid, uint32_t spin_limit_ms, hipStream_t s) {
    if (grid == 0) return (int)hipErrorInvalidValue;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;
    uint64_t spin_ticks = (uint64_t)spin_limit_ms * (uint64_t)khz;
    void *args[] = {&g, &t, &b, &max_levels, &spin_ticks};
    hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_msbfs), dim3(grid), dim3(BLOCK),
                                              args, 0, s);
    if (e != hipSuccess) return (int)e;
    return launch_status();
}
