// relax.hip -- BFS/SSSP edge relaxation over split chunks, for gfx950.
//
// Replaces M_my_drelax (cudaKernels.h:285-344) and its host fixpoint loop
// (bfs_sssp.cu:372-468), my_drelax_no_atomic + my_updateDist
// (cudaKernels.h:129-220) and the dynamic-parallelism child grids
// (my_drelax_dynamicP / dynamicP_relax_coal, cudaKernels.h:347-377,438-521).
//
// Work unit: a wave takes a batch of up to 64 chunk ids (one per lane), reads
// each chunk's bounds and its owner's distance, and walks the concatenation of
// the batch's edge ranges 64 * UNROLL slots per round (wave_owner).  A chunk
// holds at most T edges, so a hub vertex is spread over many chunks and many
// waves; this is the node split.  Chunks of one vertex are adjacent in the
// worklist and in memory, so col[]/w[] loads of a round are contiguous runs.
// Batches are dealt round-robin to the waves of the grid (no shared counter);
// the batch size shrinks with the frontier so that a small frontier still
// spreads over every wave.
//
// Relax (all modes): alt = dist[u] + w (skipped when it would wrap or reach
// inf); a plain read of dist[v] guards the atomicMin (cudaKernels.h:324-325).
// Distances only decrease, so a stale read is never lower than the truth and
// the guard cannot lose an update.  The owner's distance is read when the
// chunk is expanded, so an update made earlier in the same step is used.
//
// Next frontier, chosen per step from the size of the current one:
//   queue step (small frontier): atomicMin with return; a vertex whose
//     distance dropped is queued once per step (step stamp in mark[], taken
//     with an exchange; BFS needs none since a level-synchronous atomicMin
//     succeeds once per vertex).  Queued vertices are staged in wave-private
//     LDS and flushed as chunk ids with one atomicAdd per 64 vertices.
//   flag step (large frontier): fire-and-forget atomicMin and a plain store of
//     the step's low byte into flag[v] (a byte array: a quarter of mark[]'s
//     lines); then a three-pass compaction (count per tile, scan of tile
//     counts, write) rebuilds the list from flag[] with no shared counter.
// Device-scope atomics on MI355X execute past the per-XCD L2 and atomics on
// one address serialize (~16 ns each), so no hot path funnels through a single
// counter: batch assignment is static, work counters are striped, list
// allocation in flag steps is a scan, and the grid barrier is two-level.
//
// Bytes per unit of work (the roofline model): per edge inspected
// col 4 + w 4 (SSSP) + dist[v] 4; per chunk expanded cbeg 4 + cown 4 +
// dist[u] 4 (+ 4 for the worklist id); per relaxation 4 + 4 (queue / stamp).
#include "../gga_kernels.h"
#include "wave.hpp"
#include "sync.hpp"
#include "compact.hpp"

#include <mutex>

using namespace gga;

// Workspace of the allocation-free query (gga_sssp_noalloc): step stamps,
// flag bytes, the two chunk lists, compaction tiles, control words, work
// counters and the near-far arrays, carved by the host out of one static
// array.  Each device that runs the query holds its own copy from module
// load to process end.  Host code never names `scratch` in a kernel
// argument or a copy: it uses the device address ggak_scratch looks up.
static __device__ __align__(256) uint8_t scratch[GGA_SCRATCH_BYTES];

namespace {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t WAVES = BLOCK / GGA_WAVE;
constexpr uint32_t PBUF = 256;          // queued vertices staged per wave before a flush
#ifndef GGA_UNROLL
#define GGA_UNROLL 4
#endif
constexpr uint32_t UNROLL = GGA_UNROLL; // edge slots per lane per round in expand_batch
// Meshes: vertices a wave may follow past its batch within one step (0: off),
// its LDS queue of them, and the largest degree followed.
#ifndef GGA_CONT
#define GGA_CONT 1024
#endif
constexpr uint32_t CQ = 128;
constexpr uint32_t CONT_DEG = 16;
// Only steps whose list is at most this many chunks follow: a thin wavefront
// leaves most of the chip idle and each step is a chain of dependent loads;
// a wide one (a mesh with shortcuts, past its first levels) already fills the
// chip, and following would only expand the same vertices twice (4096^2 grid
// + 0.5% shortcuts, BFS: 5.1 ms without, 8.4 ms following in every step).
constexpr uint32_t CONT_LIST = 65536;

enum { OUT_CHANGED = 0, OUT_QUEUE = 1, OUT_FLAG = 2 };

// Work counters of one wave: wave-uniform values, kept in scalar registers
// (32 bits: one wave's share of one step), added to the striped 64-bit
// counters once per wave at the end.
struct Counters {
    uint32_t edges = 0, chunks = 0, relax = 0;
};

// A lane's chunk in a batch: its edge range [eb, ee) and owner.
struct ChunkDesc {
    uint32_t eb = 0, ee = 0, own = 0;
    bool valid = false;
};

__device__ __forceinline__ ChunkDesc load_chunk(const gga_dev_graph &g, uint32_t c, bool valid) {
    ChunkDesc d;
    d.valid = valid;
    if (valid) {
        d.eb = stream_load(&g.cbeg[c]);
        d.ee = stream_load(&g.cbeg[c + 1]);
        d.own = stream_load(&g.cown[c]);
    }
    return d;
}

// Expand one batch of chunks (this lane's chunk d when d.valid).
// thr (near-far): a destination whose new distance is >= thr is lowered but
// not listed -- it is deferred until the threshold passes it, and w.last
// (when set) records the distance each owner was expanded at, so that a
// deferred vertex is recognised by dist < last.  thr = 0xFFFFFFFF: plain
// worklist (every improvement is listed).
//
// CONT (meshes), cq given (a short list), list steps: after the batch, the wave follows
// its own wavefront -- each later round expands up to 64 vertices that this
// wave lowered below the threshold in an earlier round, from the distance it
// wrote there, all their edges (degree <= CONT_DEG), until its queue empties
// or GGA_CONT vertices were followed.  Every such vertex is listed for the
// next step all the same (stamp or flag), so the extra rounds only bring
// improvements forward: distances converge to the same fixed point.  (BFS
// then takes the stamp too: levels reached out of order can drop twice.)
template <bool BFS, bool COUNT, int OUT, bool CONT = false>
__device__ __forceinline__ void expand_batch(const gga_dev_graph &g, const gga_dev_work &w,
                                             const ChunkDesc &d, uint32_t it, uint32_t *pbuf,
                                             uint32_t &pcount, uint32_t *qout, uint32_t *tail,
                                             bool &improved, Counters &cnt, uint2 *tab, uint32_t nown,
                                             uint32_t thr = 0xFFFFFFFFu, uint32_t *cq = nullptr) {
    const uint32_t l = lane_id();
    uint32_t *__restrict__ dist = w.dist;
    const uint32_t inf = w.inf;
    uint32_t eb = 0, n = 0, du = inf;
    if (d.valid) {
        eb = d.eb;
        const uint32_t ee = d.ee;
        const uint32_t own = d.own;
        du = dist[own];
        // SSSP: an owner already lowered and listed again during this step is
        // expanded in the next one with its newer distance, which relaxes
        // every edge at least as far -- expanding it now would only repeat
        // that work, so its chunks are left to the next step.
        bool again = false;
        if (!BFS && OUT == OUT_FLAG) again = w.flag[own] == (uint8_t)it;
        if (!BFS && OUT == OUT_QUEUE) again = w.mark[own] == it;
        // Near-far: only an owner below the threshold is expanded.  Every
        // vertex listed by a stamp or a refill is; a flag byte repeats every
        // 256 steps, so a deferred vertex (dist >= thr) can also match the
        // step's byte and be listed.  Expanding it could go wrong: a
        // concurrent improvement of it would be >= thr, deferred rather than
        // stamped, and found later only through dist < last -- but with
        // several chunks reading different distances, the chunk that saw the
        // newer one may write last[] after the others, leaving dist == last
        // while edges expanded with the older distance are never relaxed
        // again.  Skipped, it is still deferred (dist < last) and a refill
        // lists it when the threshold passes it.
        const bool go = du < inf && !again && du < thr;
        n = go ? ee - eb : 0u;
        if (w.last && go) w.last[own] = du;
    }
    constexpr bool cont = CONT && OUT != OUT_CHANGED;
    uint32_t *const cqv = cq, *const cqa = cq + CQ;   // vertex, the distance this wave wrote
    uint32_t qh = 0, qt = 0, budget = (cont && cq) ? (uint32_t)GGA_CONT : 0u;   // wave-uniform
    // Lists x for the next step where `want`: the step's flag byte, or (queue
    // steps) the step stamp taken by exchange and a push.  Whole-wave call.
    auto list_vertex = [&](bool want, uint32_t x) {
        if (OUT == OUT_FLAG) {
            if (want) w.flag[x] = (uint8_t)it;
        } else if (OUT == OUT_QUEUE) {
            if ((!BFS || cont) && want) want = (w.mark[x] != it) && (atomic_xchg_u32(&w.mark[x], it) != it);
            const uint64_t mask = __ballot(want);
            if (mask) {
                const uint32_t nw = (uint32_t)__popcll(mask);
                if (pcount + nw > PBUF) {
                    flush_pushes(g, pbuf, pcount, qout, tail, w.qcap, w.ctl);
                    pcount = 0;
                }
                if (want)
                    pbuf[pcount + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u))] = x;
                pcount += nw;
                __builtin_amdgcn_wave_barrier();
            }
        }
    };
    for (;;) {
    const uint32_t incl = wave_incl_scan(n);
    const uint32_t total = wave_last(incl);
    if (COUNT) {
        cnt.edges += __builtin_amdgcn_readfirstlane(total);
        cnt.chunks += __builtin_amdgcn_readfirstlane((uint32_t)__popcll(__ballot(n > 0)));
    }
    const uint32_t excl = incl - n;
    // the wave's owner table (LDS): slot k of owner j reads edge
    // (eb_j - excl_j) + k from distance du_j -- one LDS load per slot
    // instead of two more shuffles
    tab[l] = make_uint2(eb - excl, du);
    __builtin_amdgcn_wave_barrier();
    // owners are lanes 0 .. nown - 1: the search starts at half the next
    // power of two of them (wave-uniform)
    const uint32_t half = nown <= 1u ? 0u : 1u << (31 - __builtin_clz(nown - 1u));
    // UNROLL groups of 64 slots per round.  Within a group, slots past the end
    // are clamped to the last slot, so every load below has an in-bounds
    // address and all gathers of a lane are issued before the first result is
    // used; a group wholly past the end is skipped (total and left are
    // scalars, so the skips are scalar branches) -- a small batch, e.g. one
    // chunk of a dozen edges, runs one group, not UNROLL.
    for (uint32_t base = 0; base < total; base += GGA_WAVE * UNROLL) {
        uint32_t e[UNROLL], ds[UNROLL], v[UNROLL], wt[UNROLL], cur[UNROLL], kk[UNROLL], own[UNROLL];
        bool act[UNROLL];
        const uint32_t left = total - base;   // > 0
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            const uint32_t k = base + u * GGA_WAVE + l;
            act[u] = k < total;
            kk[u] = act[u] ? k : total - 1;
            own[u] = 0;
        }
        // slot k's owner: the first lane whose inclusive prefix exceeds k
        // (binary search, one shuffle per halving; the groups' shuffles are
        // independent and issued together, skipped groups included)
        for (uint32_t step = half; step != 0; step >>= 1) {
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; u++) {
                const uint32_t y = __shfl(incl, (int)(own[u] + step - 1), GGA_WAVE);
                own[u] += y <= kk[u] ? step : 0u;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            if (u * GGA_WAVE >= left) break;
            const uint2 t = tab[own[u]];
            e[u] = t.x + kk[u];
            ds[u] = t.y;
        }
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            if (u * GGA_WAVE >= left) break;
            v[u] = stream_load(&g.col[e[u]]);
            wt[u] = BFS ? 1u : stream_load(&g.w[e[u]]);
        }
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            if (u * GGA_WAVE >= left) break;
            cur[u] = dist[v[u]];
        }
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            if (u * GGA_WAVE >= left) break;
            bool imp = false;
            const uint32_t alt = ds[u] + wt[u];
            if (act[u] && alt >= ds[u] && alt < inf && alt < cur[u]) {
                if (OUT == OUT_FLAG) {
                    (void)atomic_min_u32(&dist[v[u]], alt);  // result unused: no-return atomic
                    imp = true;
                } else {
                    const uint32_t old = atomic_min_u32(&dist[v[u]], alt);
                    imp = alt < old;
                }
            }
            if (COUNT) cnt.relax += __builtin_amdgcn_readfirstlane((uint32_t)__popcll(__ballot(imp)));
            if (OUT == OUT_CHANGED) {
                improved |= imp;
                continue;
            }
            const bool near = imp && alt < thr;
            bool queued = false;
            if (budget) {
                // queue the vertex for a later round; listed now only when the queue is full
                const uint64_t qm = __ballot(near);
                if (qm) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(qm >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)qm, 0u));
                    const uint32_t room = CQ - (qt - qh), nnear = (uint32_t)__popcll(qm);
                    queued = near && rank < room;
                    if (queued) {
                        cqv[(qt + rank) % CQ] = v[u];
                        cqa[(qt + rank) % CQ] = alt;
                    }
                    qt += nnear < room ? nnear : room;
                }
            }
            list_vertex(near && !queued, v[u]);
        }
    }
    if (qh == qt) break;
    __builtin_amdgcn_wave_barrier();   // (also: this round's table reads precede the next round's writes)
    if (budget == 0) {
        // budget spent: what is still queued is listed for the next step
        for (; qh != qt; qh += (qt - qh) < GGA_WAVE ? (qt - qh) : GGA_WAVE) {
            const bool has = l < qt - qh;
            list_vertex(has, has ? cqv[(qh + l) % CQ] : 0u);
        }
        break;
    }
    // the next round: up to 64 queued vertices, one per wave lane; a vertex
    // with more than CONT_DEG edges is listed instead.  A followed vertex
    // records the distance it was expanded at (near-far), as a listed one
    // would when the next step expands it.
    const uint32_t avail = qt - qh;
    uint32_t take = avail < GGA_WAVE ? avail : GGA_WAVE;
    take = take < budget ? take : budget;
    eb = 0;
    n = 0;
    du = inf;
    bool hub = false;
    uint32_t x = 0;
    if (l < take) {
        const uint32_t slot = (qh + l) % CQ;
        x = cqv[slot];
        du = cqa[slot];
        eb = g.cbeg[g.cstart[x]];
        n = g.cbeg[g.cstart[x + 1]] - eb;
        hub = n > CONT_DEG;
        if (hub) n = 0;
        else if (w.last) w.last[x] = du;
    }
    list_vertex(hub, x);
    qh += take;
    budget -= take;
    nown = take;
    }
}

// Work counters: lane 0 of each wave adds to one of GGA_CNT_SLOTS striped slots.
template <bool COUNT>
__device__ __forceinline__ void publish_counters(const gga_dev_work &w, const Counters &c) {
    if (COUNT && lane_id() == 0) {
        const uint32_t slot = (blockIdx.x * (blockDim.x / GGA_WAVE) + threadIdx.x / GGA_WAVE) % GGA_CNT_SLOTS;
        uint64_t *s = w.cnt + 4u * slot;
        if (c.edges) atomic_add_u64(&s[0], c.edges);
        if (c.chunks) atomic_add_u64(&s[1], c.chunks);
        if (c.relax) atomic_add_u64(&s[2], c.relax);
    }
}

// ---------------------------------------------------------------- init
__global__ __launch_bounds__(BLOCK) void k_fill(uint32_t *p, uint32_t value, uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// dist = inf except dist[src] = 0 (bfs_sssp.cu:302-318, with the source as an
// argument instead of the hard-coded parent 0); worklist = chunks of src;
// ring slot 0 = {len, step epoch}; counters and barrier zeroed.
__global__ __launch_bounds__(BLOCK) void k_query_init(gga_dev_graph g, gga_dev_work w, uint32_t src,
                                                      uint32_t epoch) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = tid; v < g.n; v += nth) w.dist[v] = (v == src) ? 0u : w.inf;
    for (uint64_t v = tid; v < g.n; v += nth) w.flag[v] = (uint8_t)(epoch - 1u);
    if (w.last)
        for (uint64_t v = tid; v < g.n; v += nth) w.last[v] = w.inf;
    for (uint64_t i = tid; i < 4ull * GGA_CNT_SLOTS; i += nth) w.cnt[i] = 0;
    if (blockIdx.x == 0) {
        const uint32_t cs = g.cstart[src], nch = g.cstart[src + 1] - cs;
        for (uint32_t i = threadIdx.x; i < nch; i += blockDim.x) w.q0[i] = cs + i;
        if (threadIdx.x < 3) {
            w.ctl[GGA_CTL_TAIL(threadIdx.x)] = threadIdx.x == 0 ? nch : 0u;
            w.ctl[GGA_CTL_ITER(threadIdx.x)] = epoch;
            // near-far: the first bucket is [0, delta); plain worklist: no bound
            w.ctl[GGA_CTL_THR(threadIdx.x)] = w.last ? (w.delta < w.inf ? w.delta : w.inf) : 0xFFFFFFFFu;
        }
        if (threadIdx.x == 0) {
            w.ctl[GGA_CTL_FAR] = 0;
            w.ctl[GGA_CTL_RSTEP] = epoch - 1u;   // no refill yet (step numbers start at epoch)
        }
        if (threadIdx.x < GGA_BAR_GROUPS) w.ctl[GGA_CTL_BARG(threadIdx.x)] = 0;
        if (threadIdx.x == 0) {
            w.ctl[GGA_CTL_CHANGED] = 0;
            w.ctl[GGA_CTL_STEPS] = 0;
            w.ctl[GGA_CTL_ERR] = 0;
            w.ctl[GGA_CTL_BAR] = 0;
        }
        if (threadIdx.x < 6) w.ctl[GGA_CTL_CNT + threadIdx.x] = 0;
        if (threadIdx.x < 8) w.ctl[GGA_CTL_RED + threadIdx.x] = 0;
    }
}

// ------------------------------------------------------ topology step
// Every chunk, every step: the reference's M_my_drelax schedule with chunks in
// place of split nodes (chunks of unreached owners contribute no edges).
template <bool BFS, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_topo_step(gga_dev_graph g, gga_dev_work w) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t c = ((uint64_t)lb * WAVES + threadIdx.x / GGA_WAVE) * GGA_WAVE + lane_id();
    const bool valid = c < g.C;
    bool improved = false;
    uint32_t pcount = 0;
    Counters cnt;
    __shared__ uint2 s_tab[WAVES][GGA_WAVE];
    expand_batch<BFS, COUNT, OUT_CHANGED>(g, w, load_chunk(g, valid ? (uint32_t)c : 0u, valid), 0u, nullptr,
                                          pcount, nullptr, nullptr, improved, cnt,
                                          s_tab[__builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE)], GGA_WAVE);
    if (__ballot(improved) && lane_id() == 0) w.ctl[GGA_CTL_CHANGED] = 1u;  // *changed = true
    publish_counters<COUNT>(w, cnt);
}

// ------------------------------------------------------ worklist step
__device__ __forceinline__ bool flag_step(const gga_dev_work &w, uint32_t nq) {
    return nq >= w.dense_min;
}

// One step = expand the list of ring slot (step % 3), fill slot (step + 1) % 3.
template <bool BFS, bool COUNT, bool CONT>
__device__ __forceinline__ void worklist_step(const gga_dev_graph &g, const gga_dev_work &w,
                                              uint32_t step, uint32_t nq, uint32_t it,
                                              uint32_t *pbuf, uint32_t *cq, uint2 *tab, Counters &cnt) {
    const uint32_t sout = (step + 1u) % 3u;
    const uint32_t *qin = (step & 1u) ? w.q1 : w.q0;
    uint32_t *qout = (step & 1u) ? w.q0 : w.q1;
    uint32_t *tail = &w.ctl[GGA_CTL_TAIL(sout)];
    // batches of bsz chunks dealt round-robin, at most 64 chunks each.
    // Ragged chunks (bmin 1): ~16 batches per wave when the list is long
    // enough, for balance; small batches also expand more chunks late in the
    // step, when more of their owners have been re-listed and can be left to
    // the next step (R-MAT s22 SSSP, batches per wave 4 / 8 / 16 / 32: 2.33 /
    // 2.17 / 2.14 / 2.26 ms).  Uniform small chunks (meshes, bmin > 1): few
    // large batches -- every batch is a chain of dependent loads -- with one
    // batch per 1/1024 of the list, but at least bmin chunks (4: each wave
    // then follows its own wavefront, see expand_batch)
    const uint32_t wpb = blockDim.x / GGA_WAVE;   // waves per workgroup (128, 256 or 512 threads)
    const uint32_t waves = gridDim.x * wpb;
    const uint32_t wid = blockIdx.x * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE);
    uint32_t bsz = w.bmin > 1u ? nq >> 10 : nq / (16u * waves);
    bsz = bsz < w.bmin ? w.bmin : bsz;
    bsz = bsz < 1u ? 1u : (bsz > GGA_WAVE ? GGA_WAVE : bsz);
    const uint32_t nbatch = (nq + bsz - 1) / bsz;
    const bool flags = flag_step(w, nq);
    const uint32_t l = lane_id();
    const uint32_t thr = w.ctl[GGA_CTL_THR(step % 3u)];
    uint32_t pcount = 0;
    bool dummy = false;
    uint32_t *const follow = (CONT && nq <= CONT_LIST) ? cq : nullptr;   // the continuation's queue
    // Software pipeline over the wave's batches b, b + waves, b + 2 waves...:
    // while a batch is expanded, the next batch's chunk descriptors and the
    // list entries of the one after it are already in flight, so a batch
    // starts at its owners' distances instead of two dependent loads
    // earlier.  The list and the chunk table do not change during a step.
    auto entry = [&](uint64_t bb, bool &ok) -> uint32_t {
        const uint64_t idx = bb * bsz + l;
        ok = bb < nbatch && l < bsz && idx < nq;
        return ok ? stream_load(&qin[idx]) : 0u;
    };
    bool ok_next;
    uint32_t c_next;
    ChunkDesc cur;
    {
        bool ok0;
        const uint32_t c0 = entry(wid, ok0);
        c_next = entry((uint64_t)wid + waves, ok_next);
        cur = load_chunk(g, c0, ok0);
    }
    for (uint64_t b = wid; b < nbatch; b += waves) {
        bool ok_after;
        const uint32_t c_after = entry(b + 2ull * waves, ok_after);
        const ChunkDesc nxt = load_chunk(g, c_next, ok_next);
        if (flags)
            expand_batch<BFS, COUNT, OUT_FLAG, CONT>(g, w, cur, it, pbuf, pcount, qout, tail, dummy, cnt, tab, bsz,
                                                     thr, follow);
        else
            expand_batch<BFS, COUNT, OUT_QUEUE, CONT>(g, w, cur, it, pbuf, pcount, qout, tail, dummy, cnt, tab, bsz,
                                                      thr, follow);
        cur = nxt;
        c_next = c_after;
        ok_next = ok_after;
    }
    if (pcount) flush_pushes(g, pbuf, pcount, qout, tail, w.qcap, w.ctl);
}

// A vertex belongs to the next list when the current (flag) step stamped it.
// One byte per vertex, so the two passes over the stamps read n bytes, not
// 4n.  The byte repeats every 256 steps of a query: a vertex last stamped 256
// steps before is listed again -- one extra expansion at an unchanged
// distance (none at all for a near-far vertex at or above the threshold,
// see expand_batch), no effect on any result.
struct StampMarked {
    const uint8_t *flag;
    uint8_t it8;
    static constexpr bool vec8 = true;
    __device__ __forceinline__ bool operator()(uint32_t v) const { return flag[v] == it8; }
    // bit i: vertex v + i holds the stamp (v a multiple of 8)
    __device__ __forceinline__ uint32_t mark8(uint64_t v) const {
        const uint2 x = *reinterpret_cast<const uint2 *>(flag + v);
        uint32_t m = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            m |= ((x.x >> (8u * i)) & 0xFFu) == it8 ? 1u << i : 0u;
            m |= ((x.y >> (8u * i)) & 0xFFu) == it8 ? 1u << (4u + i) : 0u;
        }
        return m;
    }
};

// The single list of the plain layout for step `step`.
__device__ __forceinline__ ListOut list_out(const gga_dev_graph &g, const gga_dev_work &w, uint32_t step) {
    ListOut o;
    o.cstart = g.cstart;
    o.q = (step & 1u) ? w.q0 : w.q1;
    o.qcap = w.qcap;
    o.tiles = w.tiles;
    o.tail = &w.ctl[GGA_CTL_TAIL((step + 1u) % 3u)];
    o.ctl = w.ctl;
    return o;
}

// Per-step bookkeeping done by one thread: next slot's step number, clear the
// slot consumed one step ago (it becomes the output slot of the next step).
__device__ __forceinline__ void step_bookkeeping(uint32_t *ctl, uint32_t step, uint32_t it,
                                                 uint32_t nq) {
    const uint32_t sout = (step + 1u) % 3u, sclr = (step + 2u) % 3u;
    ctl[GGA_CTL_ITER(sout)] = it + 1u;
    ctl[GGA_CTL_THR(sout)] = ctl[GGA_CTL_THR(step % 3u)];
    ctl[GGA_CTL_TAIL(sclr)] = 0u;
    ctl[GGA_CTL_LISTED(sout)] = 0u;
    if (nq) ctl[GGA_CTL_STEPS] += 1u;
}

__device__ __forceinline__ uint32_t step_len(const gga_dev_work &w, uint32_t step) {
    const uint32_t t = w.ctl[GGA_CTL_TAIL(step % 3u)];
    return t < w.qcap ? t : w.qcap;  // > qcap only with QFULL set
}

// Placement of the running wave: XCC id << 8 | SE / SH / CU bits of HW_ID.
__device__ __forceinline__ uint32_t hw_place() {
    uint32_t hw, xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    return ((xcc & 0xFu) << 8) | ((hw >> 8) & 0xFFu);
}

// TL: instrumented variant (the reference's my_drelax_metric, cudaKernels.h:2-70,
// which stores each block's clock64 span): each workgroup writes its start,
// duration and placement to tl[blockIdx.x].  Every variant keeps eight waves
// per SIMD (at most 64 VGPRs; the counting variants would otherwise take 62-65
// and drop to seven): the relax is bound by the latency of its gathers.
// BS: workgroup size, 128, 256 or 512 threads (the tuner times all three).
template <bool BFS, bool COUNT, bool TL, uint32_t BS, bool CONT>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_wl_step(gga_dev_graph g, gga_dev_work w, uint32_t step, gga_wg_record *tl) {
    __shared__ uint32_t s_pbuf[BS / GGA_WAVE][PBUF], s_cq[CONT ? BS / GGA_WAVE : 1][CONT ? 2 * CQ : 1];
    __shared__ uint2 s_tab[BS / GGA_WAVE][GGA_WAVE];
    uint64_t t0 = 0;
    if (TL && threadIdx.x == 0) t0 = wall_clock64();
    const uint32_t nq = step_len(w, step);
    const uint32_t it = w.ctl[GGA_CTL_ITER(step % 3u)];
    if (blockIdx.x == 0 && threadIdx.x == 0) step_bookkeeping(w.ctl, step, it, nq);
    if (nq != 0) {
        Counters cnt;
        const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE);
        worklist_step<BFS, COUNT, CONT>(g, w, step, nq, it, s_pbuf[wv], s_cq[CONT ? wv : 0], s_tab[wv], cnt);
        publish_counters<COUNT>(w, cnt);
    }
    if (TL) {
        __syncthreads();
        if (threadIdx.x == 0) {
            gga_wg_record r;
            r.start = t0;
            r.ticks = (uint32_t)(wall_clock64() - t0);
            r.place = hw_place();
            tl[blockIdx.x] = r;
        }
    }
}

// The compaction passes run after k_wl_step(step): count, then write (which
// also sums the counts before each tile, and with near-far buckets starts
// the refill when the list came out empty, below); both are no-ops after a
// queue step.  A count workgroup that finds stamped vertices marks the slot
// LISTED (plain stores of the same value), so that the write pass knows
// whether the list is empty without summing every count.
__global__ __launch_bounds__(BLOCK) void k_compact_count(gga_dev_graph g, gga_dev_work w, uint32_t step) {
    __shared__ uint32_t s_wave[WAVES];
    const uint32_t nq = step_len(w, step);
    if (nq == 0 || !flag_step(w, nq)) return;
    const ListOut o = list_out(g, w, step);
    const uint32_t it = w.ctl[GGA_CTL_ITER(step % 3u)];
    uint32_t any = 0;
    for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x)
        any |= compact_count_tile(o, g.n, t, StampMarked{w.flag, (uint8_t)it}, s_wave);
    if (any && threadIdx.x == 0) w.ctl[GGA_CTL_LISTED((step + 1u) % 3u)] = 1u;
}

// ---- near-far refill: after a step whose next list came out empty, list
// the deferred vertices (dist < last) under the next threshold.
struct DeferredUnder {
    const uint32_t *dist, *last;
    uint32_t t;
    uint32_t *mn = nullptr;   // when set: the smallest deferred distance seen (under t or not)
    static constexpr bool vec8 = true;
    __device__ __forceinline__ bool operator()(uint32_t v) const {
        const uint32_t d = dist[v];
        const bool deferred = d < last[v];
        if (mn && deferred && d < *mn) *mn = d;
        return deferred && d < t;
    }
    // bit i: vertex v + i is deferred under t (v a multiple of 8); dist may
    // be a caller's buffer (the allocation-free query), so 16-byte loads
    // only when it is 16-byte aligned (last always is)
    __device__ __forceinline__ uint32_t mark8(uint64_t v) const {
        uint32_t m = 0;
        if ((reinterpret_cast<uintptr_t>(dist) & 15u) == 0) {
            const uint4 d0 = *reinterpret_cast<const uint4 *>(dist + v);
            const uint4 d1 = *reinterpret_cast<const uint4 *>(dist + v + 4);
            const uint4 l0 = *reinterpret_cast<const uint4 *>(last + v);
            const uint4 l1 = *reinterpret_cast<const uint4 *>(last + v + 4);
            const uint32_t d[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
            const uint32_t s[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
            uint32_t lo = 0xFFFFFFFFu;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                m |= (d[i] < s[i] && d[i] < t) ? 1u << i : 0u;
                lo = d[i] < s[i] && d[i] < lo ? d[i] : lo;
            }
            if (mn && lo < *mn) *mn = lo;
        } else {
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) m |= (*this)((uint32_t)(v + i)) ? 1u << i : 0u;
        }
        return m;
    }
};

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const uint32_t s = a + b;
    return s < a ? 0xFFFFFFFFu : s;
}

__device__ __forceinline__ bool refill_needed(const gga_dev_work &w, uint32_t step) {
    return w.last && w.ctl[GGA_CTL_TAIL((step + 1u) % 3u)] == 0u;
}

// count: chunk ids per tile under thr + delta, and the tile's smallest deferred distance
__device__ __forceinline__ void refill_count_body(const gga_dev_graph &g, const gga_dev_work &w, uint32_t step,
                                                  uint32_t *s_wave, uint32_t *s_min) {
    const ListOut o = list_out(g, w, step);
    uint32_t m = 0xFFFFFFFFu;   // the thread's smallest deferred distance, gathered by the count
    const DeferredUnder pred{w.dist, w.last, sat_add(w.ctl[GGA_CTL_THR((step + 1u) % 3u)], w.delta), &m};
    for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x) {
        m = 0xFFFFFFFFu;
        compact_count_tile(o, g.n, t, pred, s_wave);
#pragma unroll
        for (uint32_t off = 32; off >= 1; off >>= 1) {
            const uint32_t y = __shfl_xor(m, (int)off, GGA_WAVE);
            m = y < m ? y : m;
        }
        if (lane_id() == 0) s_min[threadIdx.x / GGA_WAVE] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t x = s_min[0];
            for (uint32_t i = 1; i < WAVES; i++) x = s_min[i] < x ? s_min[i] : x;
            w.tmin[t] = x;
        }
        __syncthreads();
    }
}

// scan (one workgroup): offsets and list length; the threshold of the new
// list; when nothing lies under it but deferred vertices remain, the
// threshold jumps to the smallest deferred distance and FAR asks for another
// refill.
__device__ __forceinline__ void refill_scan_body(const gga_dev_graph &g, const gga_dev_work &w, uint32_t step,
                                                 uint32_t *s_wave, uint32_t *s_min) {
    const uint32_t sout = (step + 1u) % 3u;
    const uint32_t t1 = sat_add(w.ctl[GGA_CTL_THR(sout)], w.delta);
    const ListOut o = list_out(g, w, step);
    compact_scan(o, g.n, s_wave);
    uint32_t m = 0xFFFFFFFFu;
    for (uint32_t i = threadIdx.x; i < GGA_TILES(g.n); i += blockDim.x) m = w.tmin[i] < m ? w.tmin[i] : m;
#pragma unroll
    for (uint32_t off = 32; off >= 1; off >>= 1) {
        const uint32_t y = __shfl_xor(m, (int)off, GGA_WAVE);
        m = y < m ? y : m;
    }
    if (lane_id() == 0) s_min[threadIdx.x / GGA_WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t gmin = s_min[0];
        for (uint32_t i = 1; i < WAVES; i++) gmin = s_min[i] < gmin ? s_min[i] : gmin;
        const uint32_t total = *o.tail;
        w.ctl[GGA_CTL_RTHR] = t1;
        if (total) {
            w.ctl[GGA_CTL_THR(sout)] = t1;
            w.ctl[GGA_CTL_FAR] = 0u;
            w.ctl[GGA_CTL_RSTEP] = w.ctl[GGA_CTL_ITER(step % 3u)];
        } else if (gmin != 0xFFFFFFFFu) {
            w.ctl[GGA_CTL_THR(sout)] = gmin;   // the next refill lists [gmin, gmin + delta)
            w.ctl[GGA_CTL_FAR] = 1u;
        } else {
            w.ctl[GGA_CTL_FAR] = 0u;           // nothing deferred: converged
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void refill_write_body(const gga_dev_graph &g, const gga_dev_work &w, uint32_t step,
                                                  uint32_t *s_wave) {
    const ListOut o = list_out(g, w, step);
    const DeferredUnder pred{w.dist, w.last, w.ctl[GGA_CTL_RTHR]};
    for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x) compact_write_tile(o, g.n, t, pred, s_wave);
}

// The standalone refill is two launches: its count runs in the compaction's
// write launch (k_compact_write, whose first thread also records the
// threshold the refill lists under, RTHR, and that this step refills, RSTEP
// = its step number), then k_refill_write, which sums the counts before each
// tile itself (compact_write_tiles); its first workgroup folds the total and
// the tiles' minima into the next threshold as the scan body above does.
// That write reads only RTHR / RSTEP, which no workgroup of it changes, so
// the list length it stores midway does not change its course.
//
// In the compaction's write launch: after a flag
// step the list is empty exactly when no count workgroup marked it LISTED,
// after a queue step when no push reached it (its length, final since the
// relax ended); either way every workgroup reads the same words, which no
// workgroup of this launch changes.  (The refill's counts reuse tiles[],
// which the compaction reads only when the list is not empty.)  Its first
// thread records the threshold the refill lists under (RTHR) and that this
// step refills (RSTEP = its step number).
__global__ __launch_bounds__(BLOCK) void k_compact_write(gga_dev_graph g, gga_dev_work w, uint32_t step) {
    __shared__ uint32_t s_wave[WAVES], s_min[WAVES];
    const uint32_t nq = step_len(w, step), sout = (step + 1u) % 3u;
    const bool flags = nq != 0 && flag_step(w, nq);
    if (flags) {
        if (!w.last || w.ctl[GGA_CTL_LISTED(sout)] != 0u) {
            const uint32_t it = w.ctl[GGA_CTL_ITER(step % 3u)];
            compact_write_tiles(list_out(g, w, step), g.n, StampMarked{w.flag, (uint8_t)it}, s_wave);
            return;
        }
    } else if (!w.last || w.ctl[GGA_CTL_TAIL(sout)] != 0u) {
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.ctl[GGA_CTL_RTHR] = sat_add(w.ctl[GGA_CTL_THR(sout)], w.delta);
        w.ctl[GGA_CTL_RSTEP] = w.ctl[GGA_CTL_ITER(step % 3u)];
    }
    refill_count_body(g, w, step, s_wave, s_min);
}

__global__ __launch_bounds__(BLOCK) void k_refill_write(gga_dev_graph g, gga_dev_work w, uint32_t step) {
    __shared__ uint32_t s_wave[WAVES], s_min[WAVES];
    if (!w.last) return;
    if (w.ctl[GGA_CTL_RSTEP] != w.ctl[GGA_CTL_ITER(step % 3u)]) return;  // no refill after this step
    const uint32_t sout = (step + 1u) % 3u, t1 = w.ctl[GGA_CTL_RTHR];
    const ListOut o = list_out(g, w, step);
    if (blockIdx.x == 0) {
        // the new threshold: t1 when the refill lists anything, else a jump
        // to the smallest deferred distance (FAR: refill again), else done
        const uint32_t ntiles = GGA_TILES(g.n);
        uint32_t sum = 0, m = 0xFFFFFFFFu, ex;
        for (uint32_t i = threadIdx.x; i < ntiles; i += blockDim.x) {
            sum += o.tiles[i];
            m = w.tmin[i] < m ? w.tmin[i] : m;
        }
        const uint32_t total = block_excl_scan(sum, ex, s_wave);
#pragma unroll
        for (uint32_t off = 32; off >= 1; off >>= 1) {
            const uint32_t y = __shfl_xor(m, (int)off, GGA_WAVE);
            m = y < m ? y : m;
        }
        if (lane_id() == 0) s_min[threadIdx.x / GGA_WAVE] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t gmin = s_min[0];
            for (uint32_t i = 1; i < WAVES; i++) gmin = s_min[i] < gmin ? s_min[i] : gmin;
            if (total) {
                w.ctl[GGA_CTL_THR(sout)] = t1;
                w.ctl[GGA_CTL_FAR] = 0u;
            } else if (gmin != 0xFFFFFFFFu) {
                w.ctl[GGA_CTL_THR(sout)] = gmin;   // the next refill lists [gmin, gmin + delta)
                w.ctl[GGA_CTL_FAR] = 1u;
            } else {
                w.ctl[GGA_CTL_FAR] = 0u;           // nothing deferred: converged
            }
        }
        __syncthreads();
    }
    compact_write_tiles(o, g.n, DeferredUnder{w.dist, w.last, t1}, s_wave);
}

// ------------------------------------------------------ persistent kernel
template <bool BFS, bool COUNT, bool CONT>
__global__ __launch_bounds__(BLOCK) void k_persistent(gga_dev_graph g, gga_dev_work w,
                                                      uint32_t max_steps, uint64_t spin_ticks) {
    __shared__ uint32_t s_pbuf[WAVES][PBUF], s_cq[CONT ? WAVES : 1][CONT ? 2 * CQ : 1];
    __shared__ uint2 s_tab[WAVES][GGA_WAVE];
    __shared__ uint32_t s_state[2];
    __shared__ uint32_t s_wave[WAVES], s_min[WAVES];
    uint32_t gen = 0;
    if (threadIdx.x == 0)
        gen = __hip_atomic_load(&w.ctl[GGA_CTL_BARGEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Counters cnt;
    for (uint32_t step = 0; step < max_steps; step++) {
        const uint32_t sin = step % 3u;
        if (threadIdx.x == 0) {
            s_state[0] = __hip_atomic_load(&w.ctl[GGA_CTL_TAIL(sin)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_state[1] = __hip_atomic_load(&w.ctl[GGA_CTL_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t nq = s_state[0] < w.qcap ? s_state[0] : w.qcap, err = s_state[1];
        __syncthreads();
        if (nq == 0 || err) break;
        const uint32_t it = w.ctl[GGA_CTL_ITER(sin)];
        if (blockIdx.x == 0 && threadIdx.x == 0) step_bookkeeping(w.ctl, step, it, nq);
        const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE);
        worklist_step<BFS, COUNT, CONT>(g, w, step, nq, it, s_pbuf[wv], s_cq[CONT ? wv : 0], s_tab[wv], cnt);
        if (flag_step(w, nq)) {
            const ListOut o = list_out(g, w, step);
            grid_barrier(w.ctl, gridDim.x, gen, spin_ticks);
            for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x)
                compact_count_tile(o, g.n, t, StampMarked{w.flag, (uint8_t)it}, s_wave);
            grid_barrier(w.ctl, gridDim.x, gen, spin_ticks);
            // each workgroup sums the tile counts before its own tiles: no
            // serial scan between two more grid barriers
            compact_write_tiles(o, g.n, StampMarked{w.flag, (uint8_t)it}, s_wave);
        }
        grid_barrier(w.ctl, gridDim.x, gen, spin_ticks);
        // near-far: an empty next list is refilled from the deferred vertices,
        // as often as the threshold has to jump over empty buckets
        while (w.last) {
            if (threadIdx.x == 0) {
                s_state[0] = __hip_atomic_load(&w.ctl[GGA_CTL_TAIL((step + 1u) % 3u)], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                s_state[1] = __hip_atomic_load(&w.ctl[GGA_CTL_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            const uint32_t tail = s_state[0], rerr = s_state[1];
            __syncthreads();
            if (tail || rerr) break;
            refill_count_body(g, w, step, s_wave, s_min);
            grid_barrier(w.ctl, gridDim.x, gen, spin_ticks);
            if (blockIdx.x == 0) refill_scan_body(g, w, step, s_wave, s_min);
            grid_barrier(w.ctl, gridDim.x, gen, spin_ticks);
            if (threadIdx.x == 0) {
                s_state[0] = __hip_atomic_load(&w.ctl[GGA_CTL_TAIL((step + 1u) % 3u)], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                s_state[1] = __hip_atomic_load(&w.ctl[GGA_CTL_FAR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            const uint32_t listed = s_state[0], far = s_state[1];
            __syncthreads();
            if (listed) {
                refill_write_body(g, w, step, s_wave);
                grid_barrier(w.ctl, gridDim.x, gen, spin_ticks);
                break;
            }
            if (!far) break;   // converged: the next step finds an empty list
        }
    }
    publish_counters<COUNT>(w, cnt);
}

// ------------------------------------------------------ reductions
// Block-wide sum of one uint64 per thread (all threads get the total).
__device__ __forceinline__ uint64_t block_sum64(uint64_t x, uint64_t *s_red) {
    x = wave_sum64(x);
    if (lane_id() == 0) s_red[threadIdx.x / GGA_WAVE] = x;
    __syncthreads();
    uint64_t t = 0;
    for (uint32_t i = 0; i < WAVES; i++) t += s_red[i];
    __syncthreads();
    return t;
}

// Pass 1 (REACH_BLOCKS workgroups): per-workgroup reached / m_reach partials
// into cnt slots [0, REACH_BLOCKS) word 3 and [REACH_BLOCKS, 2 REACH_BLOCKS)
// word 3 -- plain stores, no shared counter.
constexpr uint32_t REACH_BLOCKS = GGA_CNT_SLOTS / 2;
__global__ __launch_bounds__(BLOCK) void k_reach(gga_dev_graph g, gga_dev_work w) {
    __shared__ uint64_t s_red[WAVES];
    uint64_t reached = 0, mr = 0;
    const uint64_t S = (uint64_t)gridDim.x * blockDim.x;
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // four vertices S apart per round, every load issued before the first
    // is used (a latency-bound loop otherwise: 20 us on 4.2 M vertices)
    for (; v + 3u * S < g.n; v += 4u * S) {
        uint32_t d[4], r0[4], r1[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            d[k] = w.dist[v + k * S];
            r0[k] = g.row[v + k * S];
            r1[k] = g.row[v + k * S + 1u];
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (d[k] < w.inf) {
                reached += 1;
                mr += r1[k] - r0[k];
            }
    }
    for (; v < g.n; v += S) {
        if (w.dist[v] < w.inf) {
            reached += 1;
            mr += g.row[v + 1] - g.row[v];
        }
    }
    reached = block_sum64(reached, s_red);
    mr = block_sum64(mr, s_red);
    if (threadIdx.x == 0) {
        w.cnt[4u * blockIdx.x + 3u] = reached;
        w.cnt[4u * (REACH_BLOCKS + blockIdx.x) + 3u] = mr;
    }
}

// Pass 2 (one workgroup): fold the partials into ctl[RED] and the striped work
// counters into ctl[CNT].
__global__ __launch_bounds__(BLOCK) void k_reach_fold(gga_dev_work w) {
    __shared__ uint64_t s_red[WAVES];
    uint64_t s[5] = {0, 0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < GGA_CNT_SLOTS; i += blockDim.x) {
        for (uint32_t k = 0; k < 3; k++) s[k] += w.cnt[4u * i + k];
        s[i < REACH_BLOCKS ? 3 : 4] += w.cnt[4u * i + 3u];
    }
    uint64_t *cnt = reinterpret_cast<uint64_t *>(w.ctl + GGA_CTL_CNT);
    uint64_t *red = reinterpret_cast<uint64_t *>(w.ctl + GGA_CTL_RED);
    for (uint32_t k = 0; k < 5; k++) {
        const uint64_t t = block_sum64(s[k], s_red);
        if (threadIdx.x == 0) {
            if (k < 3) cnt[k] = t;
            else red[k - 3] = t;
        }
    }
}

// Fingerprint of dist: sum over v of mix64(v << 32 | dist[v]) (the host's
// gga_checksum); partials in word 3 of cnt slots [0, REACH_BLOCKS), folded
// into ctl[RED + 4] (uint64).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(BLOCK) void k_checksum(gga_dev_graph g, gga_dev_work w) {
    __shared__ uint64_t s_red[WAVES];
    uint64_t h = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (uint64_t)gridDim.x * blockDim.x)
        h += mix64((v << 32) | w.dist[v]);
    h = block_sum64(h, s_red);
    if (threadIdx.x == 0) w.cnt[4u * blockIdx.x + 3u] = h;
}

__global__ __launch_bounds__(BLOCK) void k_checksum_fold(gga_dev_work w) {
    __shared__ uint64_t s_red[WAVES];
    uint64_t h = 0;
    for (uint32_t i = threadIdx.x; i < REACH_BLOCKS; i += blockDim.x) h += w.cnt[4u * i + 3u];
    h = block_sum64(h, s_red);
    if (threadIdx.x == 0) reinterpret_cast<uint64_t *>(w.ctl + GGA_CTL_RED)[2] = h;
}

// verifysolution (myutils.h:1-18) plus tightness, over chunks.
template <bool BFS>
__global__ __launch_bounds__(BLOCK) void k_verify_edges(gga_dev_graph g, gga_dev_work w,
                                                        uint8_t *tight) {
    const uint32_t l = lane_id();
    const uint64_t c = ((uint64_t)blockIdx.x * WAVES + threadIdx.x / GGA_WAVE) * GGA_WAVE + l;
    uint32_t eb = 0, n = 0, du = w.inf;
    if (c < g.C) {
        eb = g.cbeg[c];
        du = w.dist[g.cown[c]];
        n = du < w.inf ? g.cbeg[c + 1] - eb : 0u;
    }
    const uint32_t incl = wave_incl_scan(n);
    const uint32_t total = wave_last(incl);
    const uint32_t excl = incl - n;
    uint64_t nerr = 0;
    for (uint32_t base = 0; base < total; base += GGA_WAVE) {
        const uint32_t k = base + l;
        const uint32_t kk = k < total ? k : total - 1;
        const uint32_t j = wave_owner(incl, kk);
        const uint32_t e = wave_bcast(eb, j) + (kk - wave_bcast(excl, j));
        const uint32_t ds = wave_bcast(du, j);
        if (k < total) {
            const uint32_t v = g.col[e];
            const uint32_t wt = BFS ? 1u : g.w[e];
            const uint32_t alt = ds + wt;
            if (alt >= ds) {
                const uint32_t dv = w.dist[v];
                if (wt > 0 && alt < w.inf && alt < dv) nerr++;
                if (alt == dv) tight[v] = 1;
            }
        }
    }
    nerr = wave_sum64(nerr);
    if (l == 0 && nerr) atomic_add_u64(reinterpret_cast<uint64_t *>(w.ctl + GGA_CTL_RED), nerr);
}

__global__ __launch_bounds__(BLOCK) void k_verify_vertices(gga_dev_graph g, gga_dev_work w,
                                                           uint32_t src, const uint8_t *tight) {
    uint64_t bad = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = w.dist[v];
        if (v == src) bad += d != 0u;
        else if (d < w.inf && !tight[v]) bad++;
    }
    bad = wave_sum64(bad);
    if (lane_id() == 0 && bad) atomic_add_u64(reinterpret_cast<uint64_t *>(w.ctl + GGA_CTL_RED) + 1, bad);
}

inline int launch_status() { return (int)hipGetLastError(); }

inline uint32_t grid_for(uint64_t items, uint32_t cap) {
    uint64_t gsz = (items + BLOCK - 1) / BLOCK;
    if (gsz == 0) gsz = 1;
    return (uint32_t)(gsz < cap ? gsz : cap);
}

}  // namespace

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_addr[SCRATCH_DEVICES];   // device address of `scratch`, per device
}  // namespace

extern "C" int ggak_scratch(int device, void **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= SCRATCH_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_addr[device]) {
        // once per device, with that device current
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return (int)e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return (int)e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return (int)e;
        if (!p) return (int)hipErrorInvalidSymbol;
        scratch_addr[device] = p;
    }
    *ptr = scratch_addr[device];
    return 0;
}

extern "C" int ggak_fill_u32(uint32_t *p, uint32_t value, uint64_t count, hipStream_t s) {
    if (count == 0) return 0;
    k_fill<<<grid_for(count, 8192), BLOCK, 0, s>>>(p, value, count);
    return launch_status();
}

extern "C" int ggak_query_init(gga_dev_graph g, gga_dev_work w, uint32_t src, uint32_t epoch,
                               hipStream_t s) {
    k_query_init<<<grid_for(g.n, 8192), BLOCK, 0, s>>>(g, w, src, epoch);
    return launch_status();
}

extern "C" int ggak_topo_step(gga_dev_graph g, gga_dev_work w, int bfs, int count, hipStream_t s) {
    const uint64_t waves = ((uint64_t)g.C + GGA_WAVE - 1) / GGA_WAVE;
    const uint32_t grid = (uint32_t)((waves + WAVES - 1) / WAVES);
    if (grid == 0) return 0;
    if (bfs) {
        if (count) k_topo_step<true, true><<<grid, BLOCK, 0, s>>>(g, w);
        else k_topo_step<true, false><<<grid, BLOCK, 0, s>>>(g, w);
    } else {
        if (count) k_topo_step<false, true><<<grid, BLOCK, 0, s>>>(g, w);
        else k_topo_step<false, false><<<grid, BLOCK, 0, s>>>(g, w);
    }
    return launch_status();
}

// Kernel variants by (bfs, count, instrumented, mesh continuation) and
// workgroup size.  The continuation has kernels of its own, so that the
// registers of the others stay as they were.
template <bool BFS, bool COUNT, bool TL, bool CONT>
static const void *wl_fn_c(uint32_t block) {
    switch (block) {
    case 128: return reinterpret_cast<const void *>(&k_wl_step<BFS, COUNT, TL, 128, CONT>);
    case 512: return reinterpret_cast<const void *>(&k_wl_step<BFS, COUNT, TL, 512, CONT>);
    default: return reinterpret_cast<const void *>(&k_wl_step<BFS, COUNT, TL, 256, CONT>);
    }
}
template <bool BFS, bool COUNT, bool TL>
static const void *wl_fn(uint32_t block, int cont) {
    return cont ? wl_fn_c<BFS, COUNT, TL, true>(block) : wl_fn_c<BFS, COUNT, TL, false>(block);
}
template <bool BFS, bool COUNT>
static const void *persistent_fn(int cont) {
    return cont ? reinterpret_cast<const void *>(&k_persistent<BFS, COUNT, true>)
                : reinterpret_cast<const void *>(&k_persistent<BFS, COUNT, false>);
}
// cont: the graph is a mesh (bmin > 1) and the continuation is compiled in
static int use_cont(const gga_dev_work &w) { return GGA_CONT > 0 && w.bmin > 1u; }
static const void *pick_kernel(int persistent, int bfs, int count, uint32_t block, int cont) {
    if (persistent) {
        if (bfs) return count ? persistent_fn<true, true>(cont) : persistent_fn<true, false>(cont);
        return count ? persistent_fn<false, true>(cont) : persistent_fn<false, false>(cont);
    }
    if (bfs) return count ? wl_fn<true, true, false>(block, cont) : wl_fn<true, false, false>(block, cont);
    return count ? wl_fn<false, true, false>(block, cont) : wl_fn<false, false, false>(block, cont);
}

extern "C" int ggak_wl_step(gga_dev_graph g, gga_dev_work w, uint32_t step, int bfs, int count,
                            uint32_t grid, uint32_t block, gga_wg_record *tl, hipStream_t s) {
    if (grid == 0 || (block != 128 && block != 256 && block != 512)) return (int)hipErrorInvalidValue;
    void *args[] = {&g, &w, &step, &tl};
    // the instrumented variant always counts (its timeline goes with the work counters)
    const int cont = use_cont(w);
    const void *fn = tl ? (bfs ? wl_fn<true, true, true>(block, cont) : wl_fn<false, true, true>(block, cont))
                        : pick_kernel(0, bfs, count, block, cont);
    return (int)hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s);
}

// Convergence of one step against the final distances (printConv_rtime,
// convergence.h:100-140): diff = vertices whose distance differs; sq = sum of
// (final - cur)^2 evaluated in uint32 arithmetic as the reference does with
// its unsigned foru values (so an unreached vertex contributes
// (final - inf)^2 mod 2^32).  rows[0] += diff, rows[1] += sq.
__global__ __launch_bounds__(BLOCK) void k_conv(const uint32_t *dist, const uint32_t *fin, uint32_t n,
                                                uint64_t *rows) {
    __shared__ uint64_t s_red[WAVES];
    uint64_t diff = 0, sq = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t fs = fin[v], fi = dist[v];
        const uint32_t d = fs - fi;
        diff += fs != fi;
        sq += (uint32_t)(d * d);
    }
    diff = block_sum64(diff, s_red);
    sq = block_sum64(sq, s_red);
    if (threadIdx.x == 0) {
        if (diff) atomic_add_u64(&rows[0], diff);
        if (sq) atomic_add_u64(&rows[1], sq);
    }
}

extern "C" int ggak_conv(const uint32_t *dist, const uint32_t *fin, uint32_t n, uint64_t *rows, hipStream_t s) {
    k_conv<<<grid_for(n, 1024), BLOCK, 0, s>>>(dist, fin, n, rows);
    return launch_status();
}

extern "C" int ggak_refill(gga_dev_graph g, gga_dev_work w, uint32_t step, hipStream_t s) {
    if (!w.last) return 0;
    const uint32_t ntiles = GGA_TILES(g.n);
    const uint32_t grid = ntiles < 2048u ? (ntiles ? ntiles : 1u) : 2048u;
    k_refill_write<<<grid, BLOCK, 0, s>>>(g, w, step);   // (its count ran in k_compact_write)
    return launch_status();
}

extern "C" int ggak_compact(gga_dev_graph g, gga_dev_work w, uint32_t step, hipStream_t s) {
    const uint32_t ntiles = GGA_TILES(g.n);
    const uint32_t grid = ntiles < 2048u ? (ntiles ? ntiles : 1u) : 2048u;
    k_compact_count<<<grid, BLOCK, 0, s>>>(g, w, step);
    k_compact_write<<<grid, BLOCK, 0, s>>>(g, w, step);
    return launch_status();
}

extern "C" uint32_t ggak_wl_grid(int device, int persistent, int bfs, int count, uint32_t block) {
    int cus = 0, per = 0;
    if (persistent) block = BLOCK;
    if (block != 128 && block != 256 && block != 512) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    // (the continuation variants fit as many workgroups: same register cap,
    // LDS well within a CU's; persistent grids take one per CU either way)
    const void *fn = pick_kernel(persistent, bfs, count, block, 0);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, (int)block, 0) != hipSuccess) return 0;
    if (per <= 0) return 0;
    // A grid barrier needs every workgroup resident at once.  On gfx950 the
    // occupancy query (7-8 per CU for these kernels) overstates what the
    // hardware co-schedules (6 per CU measured for the persistent kernel), and
    // the cooperative launch does not reject the excess.  Barrier cost grows
    // with the number of workgroups (8 us at 256, 25 us at 1024 measured), so
    // the persistent grid is one workgroup per CU.
    if (persistent && per > 1) per = 1;
    return (uint32_t)cus * (uint32_t)per;
}

extern "C" int ggak_persistent(gga_dev_graph g, gga_dev_work w, int bfs, int count, uint32_t grid,
                               uint32_t max_steps, uint32_t spin_limit_ms, hipStream_t s) {
    if (grid == 0) return (int)hipErrorInvalidValue;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;  // 100 MHz
    uint64_t spin_ticks = (uint64_t)spin_limit_ms * (uint64_t)khz;
    void *args[] = {&g, &w, &max_steps, &spin_ticks};
    hipError_t e = hipLaunchCooperativeKernel(pick_kernel(1, bfs, count, BLOCK, use_cont(w)), dim3(grid),
                                              dim3(BLOCK), args, 0, s);
    if (e != hipSuccess) return (int)e;
    return launch_status();
}

extern "C" int ggak_reach(gga_dev_graph g, gga_dev_work w, hipStream_t s) {
    // REACH_BLOCKS partial slots are always all written (idle workgroups store 0)
    k_reach<<<REACH_BLOCKS, BLOCK, 0, s>>>(g, w);
    k_reach_fold<<<1, BLOCK, 0, s>>>(w);
    return launch_status();
}

extern "C" int ggak_checksum(gga_dev_graph g, gga_dev_work w, hipStream_t s) {
    k_checksum<<<REACH_BLOCKS, BLOCK, 0, s>>>(g, w);
    k_checksum_fold<<<1, BLOCK, 0, s>>>(w);
    return launch_status();
}

extern "C" int ggak_verify(gga_dev_graph g, gga_dev_work w, uint32_t src, int bfs, uint8_t *d_tight,
                           hipStream_t s) {
    const uint64_t waves = ((uint64_t)g.C + GGA_WAVE - 1) / GGA_WAVE;
    const uint32_t grid = (uint32_t)((waves + WAVES - 1) / WAVES);
    if (grid) {
        if (bfs) k_verify_edges<true><<<grid, BLOCK, 0, s>>>(g, w, d_tight);
        else k_verify_edges<false><<<grid, BLOCK, 0, s>>>(g, w, d_tight);
    }
    k_verify_vertices<<<grid_for(g.n, 4096), BLOCK, 0, s>>>(g, w, src, d_tight);
    return launch_status();
}
