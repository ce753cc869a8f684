// ms.hip -- multi-source BFS/SSSP: 32 independent traversals per step.
//
// No reference counterpart: the reference runs one source per process
// (bfs_sssp.cu:203).  Each lane is the same Bellman-Ford relaxation as the
// single-source kernels (the guard and atomicMin of M_my_drelax,
// cudaKernels.h:320-328), restricted to the lanes whose distance at the chunk
// owner dropped in the previous step; a lane converges when no vertex carries
// its bit, independently of the others, so lanes can be recycled between
// steps (the serving loop admits a new query into a finished lane).
//
// Layout: D[v * 32 + lane] (one 128-byte line per vertex).  A half wave
// relaxes one edge for all 32 lanes: lane s of the half reads dist[u][s] and
// dist[v][s] (both one coalesced line), col[e] / w[e] are broadcast loads,
// and the lanes that improved are published with one atomicOr into the
// destination's next mask.  A batch whose chunks carry few lanes each (lanes
// started far apart on a mesh rarely meet) is instead flattened into
// (edge, lane) pairs, one per wave lane, so idle lanes cost no slots.  The next list, chosen per step from the length of
// the current one: a small list pushes each vertex whose next mask was empty
// (the atomicOr's old value says which improver is first, so a vertex is
// listed once however many lanes reach it); a large list lets the compaction
// passes rebuild it from the non-zero next masks.  A finish pass then clears
// the masks the step consumed and ORs the next list's masks into its live-lane
// word.
#include "../gga_kernels.h"
#include "compact.hpp"
#include "sync.hpp"
#include "wave.hpp"

using namespace gga;

namespace {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t WAVES = BLOCK / GGA_WAVE;
constexpr uint32_t LANES = GGA_MS_LANES;
#ifndef GGA_MS_UNROLL
#define GGA_MS_UNROLL 4
#endif
constexpr uint32_t MS_UNROLL = GGA_MS_UNROLL;  // edges per half wave per round
// (edge, lane) pairs per wave lane per round of the flattened path: one
// (with the few lanes per chunk of a mesh a round rarely fills, and a deeper
// round only lengthens its chain -- 32 SSSP lanes on the 4096^2 grid: 1.28 s
// at 4, 1.20 at 2, 1.16 at 1; R-MAT s20 unchanged)
#ifndef GGA_MS_FLAT_UNROLL
#define GGA_MS_FLAT_UNROLL 1
#endif
constexpr uint32_t FLAT_UNROLL = GGA_MS_FLAT_UNROLL;
#ifndef GGA_MS_FLAT
#define GGA_MS_FLAT 16
#endif
// a batch averaging at most GGA_MS_FLAT lanes per edge slot is flattened
constexpr uint32_t PBUF = 256;                 // pushed vertices staged per wave before a flush
// Mesh engines: (vertex, lane) pairs a wave may follow past its batch within
// one step (0: off), the wave's LDS queue of them, and the largest degree
// followed (a hub is left to the next step's chunks).
#ifndef GGA_MS_CONT
#define GGA_MS_CONT 1024
#endif
#ifndef GGA_MS_CQ
#define GGA_MS_CQ 128
#endif
constexpr uint32_t CQ = GGA_MS_CQ;
constexpr uint32_t CONT_DEG = 16;

struct Counters {
    uint64_t edges = 0, chunks = 0, relax = 0;
};

template <bool COUNT>
__device__ __forceinline__ void publish_counters(const gga_dev_ms &ms, const Counters &c) {
    if (COUNT && lane_id() == 0) {
        const uint32_t slot = (blockIdx.x * WAVES + threadIdx.x / GGA_WAVE) % GGA_CNT_SLOTS;
        uint64_t *s = ms.cnt + 4u * slot;
        if (c.edges) atomic_add_u64(&s[0], c.edges);
        if (c.chunks) atomic_add_u64(&s[1], c.chunks);
        if (c.relax) atomic_add_u64(&s[2], c.relax);
    }
}

__device__ __forceinline__ const uint32_t *ms_qin(const gga_dev_ms &ms, uint32_t step) {
    return (step & 1u) ? ms.q1 : ms.q0;
}
__device__ __forceinline__ uint32_t *ms_mcur(const gga_dev_ms &ms, uint32_t step) {
    return (step & 1u) ? ms.M1 : ms.M0;
}
__device__ __forceinline__ uint32_t *ms_mnxt(const gga_dev_ms &ms, uint32_t step) {
    return (step & 1u) ? ms.M0 : ms.M1;
}
// index of (vertex v, lane s) in D for the engine's layout
__device__ __forceinline__ uint64_t didx(const gga_dev_ms &ms, uint32_t n, uint32_t v, uint32_t s) {
    return ms.lane_major ? (uint64_t)s * n + v : (uint64_t)v * LANES + s;
}

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const uint32_t r = a + b;
    return r < a ? 0xFFFFFFFFu : r;
}

__device__ __forceinline__ uint32_t ms_len(const gga_dev_ms &ms, uint32_t step) {
    const uint32_t t = ms.ctl[GGA_CTL_TAIL(step % 3u)];
    return t < ms.qcap ? t : ms.qcap;
}

struct MaskMarked {
    const uint32_t *mask;
    static constexpr bool vec8 = true;
    __device__ __forceinline__ bool operator()(uint32_t v) const { return mask[v] != 0u; }
    // bit i: vertex v + i has a non-empty mask (v a multiple of 8)
    __device__ __forceinline__ uint32_t mark8(uint64_t v) const {
        const uint4 a = *reinterpret_cast<const uint4 *>(mask + v);
        const uint4 b = *reinterpret_cast<const uint4 *>(mask + v + 4);
        return (a.x ? 1u : 0u) | (a.y ? 2u : 0u) | (a.z ? 4u : 0u) | (a.w ? 8u : 0u) |
               (b.x ? 16u : 0u) | (b.y ? 32u : 0u) | (b.z ? 64u : 0u) | (b.w ? 128u : 0u);
    }
};

__device__ __forceinline__ ListOut ms_list_out(const gga_dev_graph &g, const gga_dev_ms &ms, uint32_t step) {
    ListOut o;
    o.cstart = g.cstart;
    o.q = (step & 1u) ? ms.q0 : ms.q1;
    o.qcap = ms.qcap;
    o.tiles = ms.tiles;
    o.tail = &ms.ctl[GGA_CTL_TAIL((step + 1u) % 3u)];
    o.ctl = ms.ctl;
    return o;
}

// index of the i-th set bit of m (i < popcount(m))
__device__ __forceinline__ uint32_t nth_set_bit(uint32_t m, uint32_t i) {
    for (; i; i--) m &= m - 1u;
    return (uint32_t)__builtin_ctz(m);
}

__device__ __forceinline__ uint32_t wave_or(uint32_t x) {
#pragma unroll
    for (uint32_t o = 32; o >= 1; o >>= 1) x |= __shfl_xor(x, (int)o, GGA_WAVE);
    return x;
}

// ---------------------------------------------------------------- relax
__device__ __forceinline__ bool ms_flag_step(const gga_dev_ms &ms, uint32_t nq) { return nq >= ms.dense_min; }

// LM: lane-major layout (compile-time, so the vertex-major loop keeps its
// unconditional full-line loads)
template <bool COUNT, bool LM>
__device__ __forceinline__ void ms_relax(const gga_dev_graph &g, const gga_dev_ms &ms, uint32_t step,
                                         uint32_t nq, uint32_t *pbuf, uint32_t *nl, uint32_t *cq,
                                         Counters &cnt) {
    const bool push = !ms_flag_step(ms, nq);
    uint32_t *qout = (step & 1u) ? ms.q0 : ms.q1;
    uint32_t *tail = &ms.ctl[GGA_CTL_TAIL((step + 1u) % 3u)];
    uint32_t pcount = 0;
    const uint32_t *qin = ms_qin(ms, step);
    const uint32_t *mcur = ms_mcur(ms, step);
    uint32_t *mnxt = ms_mnxt(ms, step);
    uint32_t *__restrict__ D = ms.D;
    const uint32_t inf = ms.inf;
    const uint32_t waves = gridDim.x * WAVES;
    const uint32_t wid = blockIdx.x * WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE);
    uint32_t bsz = ms.bmin > 1u ? nq >> 10 : nq / (16u * waves);   // ~16 batches per wave on ragged graphs
    bsz = bsz < ms.bmin ? ms.bmin : bsz;
    bsz = bsz < 1u ? 1u : (bsz > GGA_WAVE ? GGA_WAVE : bsz);
    const uint32_t nbatch = (nq + bsz - 1) / bsz;
    const uint32_t l = lane_id(), half = l >> 5, s = l & (LANES - 1u);
    const uint32_t unit_word = ms.ctl[GGA_CTL_UNIT];
    const bool unit = (unit_word >> s) & 1u;
    // near-far: an improvement at or beyond the lane's threshold lowers the
    // distance but is deferred (no mask bit, a far bit instead) until a refill
    // reaches it
    const uint32_t thr_s = ms.far ? ms.ctl[GGA_CTL_LTHR(s)] : 0xFFFFFFFFu;
    const bool cont = GGA_MS_CONT > 0 && ms.bmin > 1u;
    uint32_t *const cqv = cq, *const cqs = cq + CQ, *const cqa = cq + 2u * CQ;   // vertex, lane, distance
    // lanes the wave sets in next masks, ORed into its LDS word (no register
    // stays live across the loop for it)
    if (lane_id() == 0) *nl = 0u;
    // Lists (x, lane sx) for the next step where `want`: its next-mask bit,
    // the lane word, and a push when the mask was empty.  Whole-wave call.
    auto list_pair = [&](bool want, uint32_t x, uint32_t sx) {
        bool first = false;
        if (want) {
            first = __hip_atomic_fetch_or(&mnxt[x], 1u << sx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
            __hip_atomic_fetch_or(nl, 1u << sx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
        if (push) {
            const uint64_t pm = __ballot(first);
            if (pm) {
                const uint32_t np = (uint32_t)__popcll(pm);
                if (pcount + np > PBUF) {
                    flush_pushes(g, pbuf, pcount, qout, tail, ms.qcap, ms.ctl);
                    pcount = 0;
                }
                if (first)
                    pbuf[pcount + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u))] = x;
                pcount += np;
                __builtin_amdgcn_wave_barrier();
            }
        }
    };
    for (uint32_t b = wid; b < nbatch; b += waves) {
        const uint32_t idx = b * bsz + l;
        const bool valid = l < bsz && idx < nq;
        uint32_t eb = 0, ne = 0, u = 0, m = 0;
        if (valid) {
            const uint32_t c = stream_load(&qin[idx]);
            u = stream_load(&g.cown[c]);
            m = mcur[u];
            eb = stream_load(&g.cbeg[c]);
            ne = m ? stream_load(&g.cbeg[c + 1]) - eb : 0u;
        }
        const uint32_t incl = wave_incl_scan(ne);
        const uint32_t total = wave_last(incl);
        const uint32_t excl = incl - ne;
        if (COUNT) cnt.chunks += (uint32_t)__popcll(__ballot(ne > 0));
        // (edge, lane) pairs of the batch; with few lanes per edge slot on
        // average, one pair per wave lane beats one edge per half wave
        uint32_t pc = (uint32_t)__popc(m);
        uint32_t pairs = pc * ne;
        uint32_t pincl = wave_incl_scan(pairs);
        uint32_t ptotal = wave_last(pincl);
        if (ptotal <= GGA_MS_FLAT * total) {
            // Round 0 walks the batch's owners.  On meshes (cont) each later
            // round walks up to 64 (vertex, lane) pairs that this wave
            // improved below the lane's threshold in an earlier round, from
            // the distance it wrote there: the wave follows its own wavefront
            // a few hops further within the step.  A queued pair is listed
            // for the next step only if it is not followed (queue full,
            // budget spent, degree above CONT_DEG); a followed pair had every
            // edge relaxed from the distance written, so each improvement is
            // either expanded at its value or listed, as in a plain step, and
            // the fixed point -- the result -- is unchanged.
            bool later = false;          // a continuation round (wave-uniform)
            uint32_t from = 0;           // its pair's distance at the owner
            uint32_t qh = 0, qt = 0, budget = cont ? (uint32_t)GGA_MS_CONT : 0u;
            for (;;) {
            const uint32_t pexcl = pincl - pairs;
            for (uint32_t base = 0; base < ptotal; base += GGA_WAVE * FLAT_UNROLL) {
                uint32_t v[FLAT_UNROLL], wt[FLAT_UNROLL], du[FLAT_UNROLL], dv[FLAT_UNROLL], sl[FLAT_UNROLL];
                bool act[FLAT_UNROLL];
#pragma unroll
                for (uint32_t i = 0; i < FLAT_UNROLL; i++) {
                    // pair k of owner j: edge r / pc_j, the (r % pc_j)-th lane of its mask,
                    // so consecutive pairs are the lanes of one edge (one D line)
                    const uint32_t k = base + i * GGA_WAVE + l;
                    act[i] = k < ptotal;
                    const uint32_t kk = act[i] ? k : ptotal - 1u;
                    const uint32_t j = wave_owner(pincl, kk);
                    const uint32_t r = kk - wave_bcast(pexcl, j);
                    const uint32_t pcj = wave_bcast(pc, j);
                    const uint32_t eo = r / pcj;
                    sl[i] = nth_set_bit(wave_bcast(m, j), r - eo * pcj);
                    const uint32_t e = wave_bcast(eb, j) + eo;
                    const uint32_t uu = wave_bcast(u, j);
                    v[i] = stream_load(&g.col[e]);
                    wt[i] = stream_load(&g.w[e]);
                    du[i] = later ? wave_bcast(from, j) : D[LM ? (uint64_t)sl[i] * g.n + uu : (uint64_t)uu * LANES + sl[i]];
                }
#pragma unroll
                for (uint32_t i = 0; i < FLAT_UNROLL; i++)
                    dv[i] = D[LM ? (uint64_t)sl[i] * g.n + v[i] : (uint64_t)v[i] * LANES + sl[i]];
#pragma unroll
                for (uint32_t i = 0; i < FLAT_UNROLL; i++) {
                    const uint32_t thr = (uint32_t)__shfl(thr_s, (int)sl[i], GGA_WAVE);
                    const bool on = act[i] && du[i] < inf;
                    const uint32_t a = du[i] + (((unit_word >> sl[i]) & 1u) ? 1u : wt[i]);
                    bool imp = false, first = false;
                    if (on && a >= du[i] && a < inf && a < dv[i]) {
                        const uint32_t old = atomic_min_u32(
                            &D[LM ? (uint64_t)sl[i] * g.n + v[i] : (uint64_t)v[i] * LANES + sl[i]], a);
                        imp = a < old;
                    }
                    const bool near = imp && a < thr;
                    if (imp && !near && ms.far) {
                        atomicOr(&ms.far[v[i]], 1u << sl[i]);
                        // the tile's lane summary (bits are only added during
                        // a step, so a stale read shows a subset)
                        if (!(ms.tfar[v[i] >> 6] & (1u << sl[i]))) atomicOr(&ms.tfar[v[i] >> 6], 1u << sl[i]);
                    }
                    bool queued = false;
                    if (budget) {
                        // queue the pair for a later round; listed now only when the queue is full
                        const uint64_t qm = __ballot(near);
                        if (qm) {
                            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(qm >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)qm, 0u));
                            const uint32_t room = CQ - (qt - qh), nnear = (uint32_t)__popcll(qm);
                            queued = near && rank < room;
                            if (queued) {
                                const uint32_t slot = (qt + rank) % CQ;
                                cqv[slot] = v[i];
                                cqs[slot] = sl[i];
                                cqa[slot] = a;
                            }
                            qt += nnear < room ? nnear : room;
                        }
                    }
                    if (COUNT) {
                        cnt.edges += (uint32_t)__popcll(__ballot(on));
                        cnt.relax += (uint32_t)__popcll(__ballot(imp));
                    }
                    list_pair(near && !queued, v[i], sl[i]);
                }
            }
            if (qh == qt) break;
            __builtin_amdgcn_wave_barrier();
            if (budget == 0) {
                // budget spent: what is still queued is listed for the next step
                for (; qh != qt; qh += (qt - qh) < GGA_WAVE ? (qt - qh) : GGA_WAVE) {
                    const bool has = l < qt - qh;
                    const uint32_t slot = (qh + l) % CQ;
                    list_pair(has, has ? cqv[slot] : 0u, has ? cqs[slot] : 0u);
                }
                break;
            }
            // the next round: up to 64 queued pairs, one per wave lane, each
            // owning all edges of its vertex
            const uint32_t avail = qt - qh;
            uint32_t take = avail < GGA_WAVE ? avail : GGA_WAVE;
            take = take < budget ? take : budget;
            later = true;
            u = 0; m = 0; eb = 0; from = 0;
            uint32_t ne1 = 0, sq = 0;
            bool hub = false;
            if (l < take) {
                const uint32_t slot = (qh + l) % CQ;
                u = cqv[slot];
                sq = cqs[slot];
                m = 1u << sq;
                from = cqa[slot];
                eb = g.cbeg[g.cstart[u]];
                ne1 = g.cbeg[g.cstart[u + 1]] - eb;
                hub = ne1 > CONT_DEG;
                if (hub) ne1 = 0;
                // followed from below the threshold: a far bit of the lane
                // here is stale (no deferral below the threshold can follow it)
                else if (ms.far && (ms.far[u] & m)) atomicAnd(&ms.far[u], ~m);
            }
            list_pair(hub, u, sq);   // too many edges to follow: the next step expands it
            qh += take;
            budget -= take;
            pc = l < take ? 1u : 0u;
            pairs = pc * ne1;
            pincl = wave_incl_scan(pairs);
            ptotal = wave_last(pincl);
            }
            continue;
        }
        // edge slot k goes to half (k & 1); slots past the end are clamped to
        // the last slot so every address below is in bounds
        for (uint32_t base = 0; base < total; base += 2u * MS_UNROLL) {
            uint32_t v[MS_UNROLL], wt[MS_UNROLL], du[MS_UNROLL], dv[MS_UNROLL], mm[MS_UNROLL];
            bool act[MS_UNROLL];
#pragma unroll
            for (uint32_t i = 0; i < MS_UNROLL; i++) {
                const uint32_t k = base + 2u * i + half;
                act[i] = k < total;
                const uint32_t kk = act[i] ? k : total - 1u;
                const uint32_t j = wave_owner(incl, kk);
                const uint32_t e = wave_bcast(eb, j) + (kk - wave_bcast(excl, j));
                const uint32_t uu = wave_bcast(u, j);
                mm[i] = wave_bcast(m, j);
                v[i] = stream_load(&g.col[e]);
                wt[i] = stream_load(&g.w[e]);
                // vertex-major: the whole line comes in anyway, every lane loads;
                // lane-major: only the lanes to relax gather their own word
                if (LM) du[i] = ((mm[i] >> s) & 1u) ? D[(uint64_t)s * g.n + uu] : inf;
                else du[i] = D[(uint64_t)uu * LANES + s];
            }
#pragma unroll
            for (uint32_t i = 0; i < MS_UNROLL; i++) {
                if (LM) dv[i] = ((mm[i] >> s) & 1u) ? D[(uint64_t)s * g.n + v[i]] : 0u;
                else dv[i] = D[(uint64_t)v[i] * LANES + s];
            }
#pragma unroll
            for (uint32_t i = 0; i < MS_UNROLL; i++) {
                const bool on = act[i] && ((mm[i] >> s) & 1u) && du[i] < inf;
                const uint32_t a = du[i] + (unit ? 1u : wt[i]);
                bool imp = false;
                if (on && a >= du[i] && a < inf && a < dv[i]) {
                    const uint32_t old =
                        atomic_min_u32(&D[LM ? (uint64_t)s * g.n + v[i] : (uint64_t)v[i] * LANES + s], a);
                    imp = a < old;
                }
                const uint64_t bal = __ballot(imp && a < thr_s);
                const uint32_t hm = (uint32_t)(bal >> (32u * half));
                bool first = false;
                if (s == 0 && hm) {
                    first = __hip_atomic_fetch_or(&mnxt[v[i]], hm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
                    __hip_atomic_fetch_or(nl, hm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                }
                if (ms.far) {
                    const uint32_t fm = (uint32_t)(__ballot(imp && a >= thr_s) >> (32u * half));
                    if (s == 0 && fm) {
                        atomicOr(&ms.far[v[i]], fm);
                        if ((ms.tfar[v[i] >> 6] & fm) != fm) atomicOr(&ms.tfar[v[i] >> 6], fm);
                    }
                }
                if (COUNT) {
                    cnt.edges += (uint32_t)__popcll(__ballot(on));
                    cnt.relax += (uint32_t)__popcll(__ballot(imp));
                }
                if (push) {
                    // the improver that found the next mask empty lists the vertex
                    const uint64_t pm = __ballot(first);
                    if (pm) {
                        const uint32_t np = (uint32_t)__popcll(pm);
                        if (pcount + np > PBUF) {
                            flush_pushes(g, pbuf, pcount, qout, tail, ms.qcap, ms.ctl);
                            pcount = 0;
                        }
                        if (first) {
                            const uint32_t pos = pcount + __builtin_amdgcn_mbcnt_hi(
                                                              (uint32_t)(pm >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
                            pbuf[pos] = v[i];
                        }
                        pcount += np;
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
        }
    }
    if (pcount) flush_pushes(g, pbuf, pcount, qout, tail, ms.qcap, ms.ctl);
    // the next list's lane word: every lane with a bit in a next mask
    __builtin_amdgcn_wave_barrier();
    if (lane_id() == 0) {
        const uint32_t x = __hip_atomic_load(nl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        if (x)
            __hip_atomic_fetch_or(&ms.ctl[GGA_CTL_LANES((step + 1u) % 3u)], x, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ------------------------------------------------------- finish pass
// Clears the masks step `step` consumed (owners of its list) and the slot the
// step after next fills (its length and lanes), which held the list of
// step - 1.  (The relax itself ORs the next list's lanes into its slot.)
__device__ __forceinline__ void ms_finish(const gga_dev_graph &g, const gga_dev_ms &ms, uint32_t step) {
    const uint32_t nin = ms_len(ms, step);
    const uint32_t *qin = ms_qin(ms, step);
    uint32_t *mcur = ms_mcur(ms, step);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // the thread holding an owner's first chunk retires the owner: clears its
    // mask and (near-far) the far bits of the lanes it expanded.  No expanded
    // lane can have been deferred at the owner during the step: its distance
    // was under the threshold, so every drop during the step was too.  (The
    // far word is cleared with an atomic: a refill may update it meanwhile.)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nin; i += stride) {
        const uint32_t c = qin[i];
        const uint32_t u = g.cown[c];
        if (c != g.cstart[u]) continue;
        const uint32_t m = mcur[u];
        if (ms.far && (ms.far[u] & m)) atomicAnd(&ms.far[u], ~m);
        mcur[u] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ms.ctl[GGA_CTL_TAIL((step + 2u) % 3u)] = 0u;
        ms.ctl[GGA_CTL_LANES((step + 2u) % 3u)] = 0u;
        if (nin) ms.ctl[GGA_CTL_STEPS] += 1u;
    }
}

// ---------------------------------------------------------------- kernels
// Lanes in `lanes` restart: dist = inf except 0 at the lane's source.
__global__ __launch_bounds__(BLOCK) void k_ms_reset(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes,
                                                    gga_ms_sources src) {
    const uint64_t total = (uint64_t)g.n * LANES;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)(i & (LANES - 1u)), v = (uint32_t)(i / LANES);
        if ((lanes >> s) & 1u) ms.D[i] = v == src.v[s] ? 0u : ms.inf;
        if (s == 0 && ms.far) {
            ms.far[v] &= ~lanes;
            if ((v & 63u) == 0) ms.tfar[v >> 6] &= ~lanes;
        }
    }
}

// Lane-major form: blockIdx.y = lane, coalesced over vertices.
__global__ __launch_bounds__(BLOCK) void k_ms_reset_lm(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes,
                                                       gga_ms_sources src) {
    const uint32_t s = blockIdx.y;
    if (!((lanes >> s) & 1u)) return;
    uint32_t *row = ms.D + (uint64_t)s * g.n;
    // the row of the lowest restarted lane clears the far bits of all of them
    const bool clear_far = ms.far && s == (uint32_t)__builtin_ctz(lanes);
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (uint64_t)gridDim.x * blockDim.x) {
        row[v] = v == src.v[s] ? 0u : ms.inf;
        if (clear_far) {
            ms.far[v] &= ~lanes;
            if ((v & 63u) == 0) ms.tfar[v >> 6] &= ~lanes;
        }
    }
}

// Seeds the restarted lanes into the list of ring slot `step`: the source's
// mask gets the lane bit; the first lane to reach an unlisted source appends
// its chunks.
__global__ __launch_bounds__(64) void k_ms_seed(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, uint32_t unit,
                                                gga_ms_sources src, uint32_t step) {
    const uint32_t s = threadIdx.x;
    if (s == 0) {
        ms.ctl[GGA_CTL_UNIT] = (ms.ctl[GGA_CTL_UNIT] & ~lanes) | (unit & lanes);
        ms.ctl[GGA_CTL_RUN] |= lanes;
    }
    if (s < LANES && ((lanes >> s) & 1u))   // BFS lanes are one bucket per level already
        ms.ctl[GGA_CTL_LTHR(s)] = (ms.far && ms.delta && !((unit >> s) & 1u)) ? ms.delta : 0xFFFFFFFFu;
    if (s >= LANES || !((lanes >> s) & 1u)) return;
    const uint32_t x = src.v[s];
    uint32_t *mcur = ms_mcur(ms, step);
    const uint32_t old = __hip_atomic_fetch_or(&mcur[x], 1u << s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(&ms.ctl[GGA_CTL_LANES(step % 3u)], 1u << s, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    if (old != 0u) return;  // already listed
    const uint32_t cs = g.cstart[x], nch = g.cstart[x + 1] - cs;
    const uint32_t pos = atomic_add_u32(&ms.ctl[GGA_CTL_TAIL(step % 3u)], nch);
    uint32_t *q = (step & 1u) ? ms.q1 : ms.q0;
    for (uint32_t k = 0; k < nch; k++) queue_store(q, pos + k, cs + k, ms.qcap, ms.ctl);
}

template <bool COUNT, bool LM>
// Eight waves per SIMD (64 VGPRs): the compiler's own choice, 67 VGPRs, allows
// seven; the eighth wave of memory-bound relax took 32 SSSP lanes on R-MAT s20
// from 7.4 to 5.9 ms (grids unchanged).
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_ms_step(gga_dev_graph g, gga_dev_ms ms, uint32_t step) {
    __shared__ uint32_t s_pbuf[WAVES][PBUF], s_nl[WAVES], s_cq[WAVES][3 * CQ];
    const uint32_t nq = ms_len(ms, step);
    if (nq == 0) return;
    Counters cnt;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE);   // wave-uniform
    ms_relax<COUNT, LM>(g, ms, step, nq, s_pbuf[wv], &s_nl[wv], s_cq[wv], cnt);
    publish_counters<COUNT>(ms, cnt);
}

// The compaction passes after a flag step, count then write (which sums the
// counts before each tile itself); no-ops after a push step.
__global__ __launch_bounds__(CBLOCK) void k_ms_count(gga_dev_graph g, gga_dev_ms ms, uint32_t step) {
    __shared__ uint32_t s_wave[CBLOCK / GGA_WAVE];
    const uint32_t nq = ms_len(ms, step);
    if (nq == 0 || !ms_flag_step(ms, nq)) return;
    const ListOut o = ms_list_out(g, ms, step);
    const MaskMarked mk{ms_mnxt(ms, step)};
    for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x) compact_count_tile(o, g.n, t, mk, s_wave);
}

__global__ __launch_bounds__(CBLOCK) void k_ms_write(gga_dev_graph g, gga_dev_ms ms, uint32_t step) {
    __shared__ uint32_t s_wave[CBLOCK / GGA_WAVE];
    const uint32_t nq = ms_len(ms, step);
    if (nq == 0 || !ms_flag_step(ms, nq)) return;
    compact_write_tiles(ms_list_out(g, ms, step), g.n, MaskMarked{ms_mnxt(ms, step)}, s_wave);
}

// ---- the tail of a step: the finish pass, then (near-far) the refill of
// lanes still running but without listed work (candidates): they list their
// deferred (vertex, lane) pairs -- the far bits -- whose distance is under
// threshold + delta, and clear those bits; a vertex whose next mask was empty
// is appended to the next list (wave-staged pushes, as in a push step).
// Per-lane minima of the deferred distances beyond it go to lpart (one row of
// 32 per workgroup); the last workgroup to finish reduces them (a counter,
// no waiting): a lane whose refill listed pairs raises its threshold by
// delta; with nothing under the new threshold but deferred pairs left, the
// threshold jumps to the smallest deferred distance (listed by the next
// step's refill); with nothing deferred the lane has converged.  The scan
// reads one far-tile word per 64 vertices (the lanes deferred in the tile),
// the far word of each vertex in a tile with a candidate's lane, and D only
// where a candidate's bit is set.
constexpr uint32_t RBLOCKS = GGA_MS_RBLOCKS;
__global__ __launch_bounds__(BLOCK) void k_ms_tail(gga_dev_graph g, gga_dev_ms ms, uint32_t step) {
    __shared__ uint32_t s_t1[LANES], s_min[LANES], s_listed, s_last, s_done;
    __shared__ uint32_t s_pbuf[WAVES][PBUF];
    __shared__ uint32_t s_red[BLOCK / LANES][LANES];
    ms_finish(g, ms, step);
    if (!ms.far) return;
    const uint32_t sout = (step + 1u) % 3u;
    // RUN and LTHR change only in the last workgroup's reduction, after every
    // workgroup has read them
    const uint32_t cand = ms.ctl[GGA_CTL_RUN] & ~ms.ctl[GGA_CTL_LANES(sout)];
    if (!cand) return;
    uint32_t *mnxt = ms_mnxt(ms, step);
    uint32_t *qout = (step & 1u) ? ms.q0 : ms.q1;
    uint32_t *tail = &ms.ctl[GGA_CTL_TAIL(sout)];
    uint32_t *pbuf = s_pbuf[threadIdx.x / GGA_WAVE];
    if (threadIdx.x < LANES) {
        s_t1[threadIdx.x] = sat_add(ms.ctl[GGA_CTL_LTHR(threadIdx.x)], ms.delta);
        s_min[threadIdx.x] = 0xFFFFFFFFu;
    }
    if (threadIdx.x == 0) s_listed = 0;
    __syncthreads();
    uint32_t listed = 0, pcount = 0;
    // Each wave reads 64 consecutive far-tile words in one load, then walks
    // only the tiles holding a candidate lane's bits, one vertex per lane
    // (wave-uniform: the pushes below use whole-wave operations).
    const uint64_t ntiles = GGA_MS_FAR_TILES(g.n);
    const uint64_t wstride = (uint64_t)gridDim.x * WAVES * GGA_WAVE;
    for (uint64_t tb = ((uint64_t)blockIdx.x * WAVES + threadIdx.x / GGA_WAVE) * GGA_WAVE; tb < ntiles;
         tb += wstride) {
        const uint64_t tl = tb + lane_id();
        uint64_t tm = __ballot(tl < ntiles && (ms.tfar[tl] & cand) != 0u);   // tiles with a candidate's bit
        while (tm) {
            const uint64_t tile = tb + (uint64_t)__builtin_ctzll(tm);
            tm &= tm - 1u;
            const uint64_t v = tile * GGA_WAVE + lane_id();
            bool first = false;
            uint32_t left = 0;   // far bits this lane keeps
            if (v < g.n) {
                const uint32_t fw = ms.far[v];
                left = fw;
                uint32_t m = fw & cand, bits = 0;
                while (m) {
                    const uint32_t s = (uint32_t)__builtin_ctz(m);
                    m &= m - 1u;
                    const uint32_t d = ms.D[didx(ms, g.n, (uint32_t)v, s)];
                    if (d < s_t1[s]) bits |= 1u << s;
                    else atomicMin(&s_min[s], d);
                }
                if (bits) {
                    left = atomicAnd(&ms.far[v], ~bits) & ~bits;
                    listed |= bits;
                    first = __hip_atomic_fetch_or(&mnxt[v], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
                }
            }
            // the lanes still deferred in the tile (far bits are only cleared
            // while this kernel runs; the step that sets them ran before, and
            // each tile has one wave here)
            const uint32_t keep = wave_or(left);
            if (lane_id() == 0) ms.tfar[tile] = keep;
            const uint64_t pm = __ballot(first);
            if (pm) {
                const uint32_t np = (uint32_t)__popcll(pm);
                if (pcount + np > PBUF) {
                    flush_pushes(g, pbuf, pcount, qout, tail, ms.qcap, ms.ctl);
                    pcount = 0;
                }
                if (first)
                    pbuf[pcount + __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u))] = (uint32_t)v;
                pcount += np;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (pcount) flush_pushes(g, pbuf, pcount, qout, tail, ms.qcap, ms.ctl);
    listed = wave_or(listed);
    if (lane_id() == 0 && listed) atomicOr(&s_listed, listed);
    __syncthreads();
    if (threadIdx.x == 0) {
        // the arriving thread writes the workgroup's lpart row itself, so
        // the release of its arrival covers the row (rather than an
        // agent-scope fence in every thread of every workgroup)
        for (uint32_t s = 0; s < LANES; s++) ms.lpart[(uint64_t)blockIdx.x * LANES + s] = s_min[s];
        if (s_listed)
            __hip_atomic_fetch_or(&ms.ctl[GGA_CTL_RLISTED], s_listed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = __hip_atomic_fetch_add(&ms.ctl[GGA_CTL_TDONE], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                 gridDim.x - 1u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();   // every row, as its workgroup left it
    {
        const uint32_t s = threadIdx.x & (LANES - 1u), grp = threadIdx.x / LANES;
        uint32_t gmin = 0xFFFFFFFFu;
        for (uint32_t b = grp; b < gridDim.x; b += BLOCK / LANES) {
            const uint32_t x = ms.lpart[(uint64_t)b * LANES + s];
            gmin = x < gmin ? x : gmin;
        }
        s_red[grp][s] = gmin;
    }
    if (threadIdx.x == 0) s_done = 0;
    __syncthreads();
    const uint32_t rl = __hip_atomic_load(&ms.ctl[GGA_CTL_RLISTED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < LANES && ((cand >> threadIdx.x) & 1u)) {
        const uint32_t s = threadIdx.x;
        uint32_t gmin = 0xFFFFFFFFu;
        for (uint32_t k = 0; k < BLOCK / LANES; k++) gmin = s_red[k][s] < gmin ? s_red[k][s] : gmin;
        if ((rl >> s) & 1u) ms.ctl[GGA_CTL_LTHR(s)] = sat_add(ms.ctl[GGA_CTL_LTHR(s)], ms.delta);
        else if (gmin != 0xFFFFFFFFu) ms.ctl[GGA_CTL_LTHR(s)] = gmin;
        else atomicOr(&s_done, 1u << s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ms.ctl[GGA_CTL_RUN] &= ~s_done;
        ms.ctl[GGA_CTL_LANES(sout)] |= rl;
        ms.ctl[GGA_CTL_RLISTED] = 0u;
        ms.ctl[GGA_CTL_TDONE] = 0u;
    }
}

// All steps in one cooperative launch: relax, then the compaction's count
// and write passes, separated by grid barriers.  Step numbers continue from `step0`.
template <bool COUNT, bool LM>
__global__ __launch_bounds__(BLOCK) void k_ms_persistent(gga_dev_graph g, gga_dev_ms ms, uint32_t step0,
                                                         uint32_t max_steps, uint64_t spin_ticks) {
    __shared__ uint32_t s_pbuf[WAVES][PBUF];
    __shared__ uint32_t s_wave[WAVES], s_state[2], s_nl[WAVES], s_cq[WAVES][3 * CQ];
    uint32_t gen = 0;
    if (threadIdx.x == 0)
        gen = __hip_atomic_load(&ms.ctl[GGA_CTL_BARGEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Counters cnt;
    for (uint32_t k = 0; k < max_steps; k++) {
        const uint32_t step = step0 + k;
        if (threadIdx.x == 0) {
            s_state[0] = __hip_atomic_load(&ms.ctl[GGA_CTL_TAIL(step % 3u)], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            s_state[1] = __hip_atomic_load(&ms.ctl[GGA_CTL_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t nq = s_state[0] < ms.qcap ? s_state[0] : ms.qcap, err = s_state[1];
        __syncthreads();
        if (nq == 0 || err) break;
        const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x / GGA_WAVE);
        ms_relax<COUNT, LM>(g, ms, step, nq, s_pbuf[wv], &s_nl[wv], s_cq[wv], cnt);
        grid_barrier(ms.ctl, gridDim.x, gen, spin_ticks);
        if (ms_flag_step(ms, nq)) {
            const ListOut o = ms_list_out(g, ms, step);
            const MaskMarked mk{ms_mnxt(ms, step)};
            for (uint32_t t = blockIdx.x; t < GGA_TILES(g.n); t += gridDim.x) compact_count_tile(o, g.n, t, mk, s_wave);
            grid_barrier(ms.ctl, gridDim.x, gen, spin_ticks);
            // each workgroup sums the tile counts before its own tiles
            compact_write_tiles(o, g.n, mk, s_wave);
            grid_barrier(ms.ctl, gridDim.x, gen, spin_ticks);
        }
        ms_finish(g, ms, step);
        grid_barrier(ms.ctl, gridDim.x, gen, spin_ticks);
    }
    publish_counters<COUNT>(ms, cnt);
}

// Lane results out of the vertex-major table: a workgroup transposes 64
// vertices x 32 lanes through LDS and writes each selected lane's 64 values as
// one contiguous run of out.p[lane].
constexpr uint32_t HV = 64;
__global__ __launch_bounds__(BLOCK) void k_ms_harvest(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes,
                                                      gga_ms_outputs out) {
    __shared__ uint32_t tile[HV][LANES + 1];
    const uint64_t v0 = (uint64_t)blockIdx.x * HV;
    for (uint32_t i = threadIdx.x; i < HV * LANES; i += blockDim.x) {
        const uint64_t v = v0 + i / LANES;
        tile[i / LANES][i % LANES] = v < g.n ? ms.D[v * LANES + (i % LANES)] : 0u;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < HV * LANES; i += blockDim.x) {
        const uint32_t s = i / HV, j = i % HV;
        const uint64_t v = v0 + j;
        if (((lanes >> s) & 1u) && v < g.n) out.p[s][v] = tile[j][s];
    }
}

__global__ __launch_bounds__(BLOCK) void k_ms_harvest_lm(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes,
                                                         gga_ms_outputs out) {
    const uint32_t s = blockIdx.y;
    if (!((lanes >> s) & 1u)) return;
    const uint32_t *row = ms.D + (uint64_t)s * g.n;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (uint64_t)gridDim.x * blockDim.x)
        out.p[s][v] = row[v];
}

// Reached (vertex, lane) pairs of `lanes` and the out-degree sum over them
// (the GTEPS numerator): per-workgroup partials in word 3 of cnt slots
// [0, REACH) and [REACH, 2 REACH), folded by k_ms_fold.
constexpr uint32_t REACH = GGA_CNT_SLOTS / 2;
__global__ __launch_bounds__(BLOCK) void k_ms_reach(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes) {
    __shared__ uint64_t s_red[2][WAVES];
    uint64_t r = 0, mr = 0;
    const uint32_t s = threadIdx.x & (LANES - 1u);
    const bool mine = (lanes >> s) & 1u;
    // thread i of the grid reads D[i]: lane i % 32 of vertex i / 32 (coalesced);
    // four entries S apart per round, every load issued before the first is
    // used (a latency-bound loop otherwise: 145 us for 32 lanes of 1 M vertices)
    const uint64_t total = (uint64_t)g.n * LANES, S = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3u * S < total; i += 4u * S) {
        uint32_t d[4], r0[4], r1[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint64_t v = (i + k * S) / LANES;
            d[k] = ms.D[i + k * S];
            r0[k] = g.row[v];
            r1[k] = g.row[v + 1];
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
            if (mine && d[k] < ms.inf) {
                r += 1;
                mr += r1[k] - r0[k];
            }
    }
    for (; i < total; i += S) {
        if (mine && ms.D[i] < ms.inf) {
            const uint32_t v = (uint32_t)(i / LANES);
            r += 1;
            mr += g.row[v + 1] - g.row[v];
        }
    }
    r = wave_sum64(r);
    mr = wave_sum64(mr);
    if (lane_id() == 0) {
        s_red[0][threadIdx.x / GGA_WAVE] = r;
        s_red[1][threadIdx.x / GGA_WAVE] = mr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0, b = 0;
        for (uint32_t i = 0; i < WAVES; i++) {
            a += s_red[0][i];
            b += s_red[1][i];
        }
        ms.cnt[4u * blockIdx.x + 3u] = a;
        ms.cnt[4u * (REACH + blockIdx.x) + 3u] = b;
    }
}

// Lane-major reach: a 16 x 32 grid, one partial slot per workgroup.
__global__ __launch_bounds__(BLOCK) void k_ms_reach_lm(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes) {
    __shared__ uint64_t s_red[2][WAVES];
    const uint32_t s = blockIdx.y;
    uint64_t r = 0, mr = 0;
    if ((lanes >> s) & 1u) {
        const uint32_t *row = ms.D + (uint64_t)s * g.n;
        const uint64_t S = (uint64_t)gridDim.x * blockDim.x;
        uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        for (; v + 3u * S < g.n; v += 4u * S) {   // four vertices in flight, as above
            uint32_t d[4], r0[4], r1[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                d[k] = row[v + k * S];
                r0[k] = g.row[v + k * S];
                r1[k] = g.row[v + k * S + 1u];
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; k++)
                if (d[k] < ms.inf) {
                    r += 1;
                    mr += r1[k] - r0[k];
                }
        }
        for (; v < g.n; v += S)
            if (row[v] < ms.inf) {
                r += 1;
                mr += g.row[v + 1] - g.row[v];
            }
    }
    r = wave_sum64(r);
    mr = wave_sum64(mr);
    if (lane_id() == 0) {
        s_red[0][threadIdx.x / GGA_WAVE] = r;
        s_red[1][threadIdx.x / GGA_WAVE] = mr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0, b = 0;
        for (uint32_t i = 0; i < WAVES; i++) {
            a += s_red[0][i];
            b += s_red[1][i];
        }
        const uint32_t slot = blockIdx.y * gridDim.x + blockIdx.x;
        ms.cnt[4u * slot + 3u] = a;
        ms.cnt[4u * (REACH + slot) + 3u] = b;
    }
}

// Per-lane fingerprint of the distances: sum over v of mix64(v << 32 | D[v][lane])
// (wrapping; order-independent), the value gga_checksum computes on the host.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(BLOCK) void k_ms_checksum(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes,
                                                       uint64_t *sums) {
    __shared__ uint64_t s_part[BLOCK];
    const uint32_t s = threadIdx.x & (LANES - 1u);
    const bool mine = (lanes >> s) & 1u;
    uint64_t h = 0;
    const uint64_t total = (uint64_t)g.n * LANES;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (mine) h += mix64(((i / LANES) << 32) | ms.D[i]);
    s_part[threadIdx.x] = h;
    __syncthreads();
    if (threadIdx.x < LANES && mine) {
        uint64_t t = 0;
        for (uint32_t k = threadIdx.x; k < BLOCK; k += LANES) t += s_part[k];
        atomic_add_u64(&sums[threadIdx.x], t);
    }
}

__global__ __launch_bounds__(BLOCK) void k_ms_checksum_lm(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes,
                                                          uint64_t *sums) {
    __shared__ uint64_t s_red[WAVES];
    const uint32_t s = blockIdx.y;
    if (!((lanes >> s) & 1u)) return;
    const uint32_t *row = ms.D + (uint64_t)s * g.n;
    uint64_t h = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += (uint64_t)gridDim.x * blockDim.x)
        h += mix64((v << 32) | row[v]);
    h = wave_sum64(h);
    if (lane_id() == 0) s_red[threadIdx.x / GGA_WAVE] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t i = 0; i < WAVES; i++) t += s_red[i];
        atomic_add_u64(&sums[s], t);
    }
}

// Work counters: fold the striped slots into ctl[CNT] (one workgroup) and the
// reach partials into ctl[RED] (2 x uint64).
__global__ __launch_bounds__(BLOCK) void k_ms_fold(gga_dev_ms ms) {
    __shared__ uint64_t s_red[5][WAVES];
    uint64_t x[5] = {0, 0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < GGA_CNT_SLOTS; i += blockDim.x) {
        x[0] += ms.cnt[4u * i];
        x[1] += ms.cnt[4u * i + 1u];
        x[2] += ms.cnt[4u * i + 2u];
        x[i < REACH ? 3 : 4] += ms.cnt[4u * i + 3u];
    }
#pragma unroll
    for (uint32_t k = 0; k < 5; k++) {
        x[k] = wave_sum64(x[k]);
        if (lane_id() == 0) s_red[k][threadIdx.x / GGA_WAVE] = x[k];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        uint64_t t = 0;
        for (uint32_t i = 0; i < WAVES; i++) t += s_red[threadIdx.x][i];
        if (threadIdx.x < 3) reinterpret_cast<uint64_t *>(ms.ctl + GGA_CTL_CNT)[threadIdx.x] = t;
        else reinterpret_cast<uint64_t *>(ms.ctl + GGA_CTL_RED)[threadIdx.x - 3] = t;
    }
}

// Counters, ring slots, lane masks and barrier words start at zero.
__global__ __launch_bounds__(BLOCK) void k_ms_clear(gga_dev_ms ms) {
    for (uint32_t i = threadIdx.x; i < 4u * GGA_CNT_SLOTS; i += blockDim.x) ms.cnt[i] = 0;
    for (uint32_t i = threadIdx.x; i < GGA_CTL_WORDS; i += blockDim.x) ms.ctl[i] = 0;
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

extern "C" int ggak_ms_clear(gga_dev_ms ms, hipStream_t s) {
    k_ms_clear<<<1, BLOCK, 0, s>>>(ms);
    return launch_status();
}

extern "C" int ggak_ms_start(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, uint32_t unit, gga_ms_sources src,
                             uint32_t step, hipStream_t s) {
    if (lanes == 0) return 0;
    for (uint32_t i = 0; i < LANES; i++)
        if (((lanes >> i) & 1u) && src.v[i] >= g.n) return (int)hipErrorInvalidValue;
    if (ms.lane_major) k_ms_reset_lm<<<dim3(grid_for(g.n, 256), LANES), BLOCK, 0, s>>>(g, ms, lanes, src);
    else k_ms_reset<<<grid_for((uint64_t)g.n * LANES, 8192), BLOCK, 0, s>>>(g, ms, lanes, src);
    k_ms_seed<<<1, 64, 0, s>>>(g, ms, lanes, unit, src, step);
    return launch_status();
}

template <bool COUNT, bool LM>
static const void *step_fn() {
    return reinterpret_cast<const void *>(&k_ms_step<COUNT, LM>);
}
template <bool COUNT, bool LM>
static const void *persistent_fn() {
    return reinterpret_cast<const void *>(&k_ms_persistent<COUNT, LM>);
}
static const void *pick(int persistent, int count, uint32_t lm) {
    static const void *const st[4] = {step_fn<false, false>(), step_fn<false, true>(), step_fn<true, false>(),
                                      step_fn<true, true>()};
    static const void *const pk[4] = {persistent_fn<false, false>(), persistent_fn<false, true>(),
                                      persistent_fn<true, false>(), persistent_fn<true, true>()};
    const int i = (count ? 2 : 0) | (lm ? 1 : 0);
    return persistent ? pk[i] : st[i];
}

extern "C" int ggak_ms_step(gga_dev_graph g, gga_dev_ms ms, uint32_t step, int count, uint32_t grid,
                            hipStream_t s) {
    if (grid == 0) return (int)hipErrorInvalidValue;
    void *args[] = {&g, &ms, &step};
    hipError_t e = hipLaunchKernel(pick(0, count, ms.lane_major), dim3(grid), dim3(BLOCK), args, 0, s);
    if (e != hipSuccess) return (int)e;
    const uint32_t cg = compact_grid(g.n);
    if (ms.dense_min != 0xFFFFFFFFu) {   // push-only engines never rebuild from the masks
        k_ms_count<<<cg, CBLOCK, 0, s>>>(g, ms, step);
        k_ms_write<<<cg, CBLOCK, 0, s>>>(g, ms, step);
    }
    // lpart holds one row per workgroup of the tail pass
    k_ms_tail<<<RBLOCKS, BLOCK, 0, s>>>(g, ms, step);
    return launch_status();
}

extern "C" uint32_t ggak_ms_grid(int device, int persistent, int count) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, pick(persistent, count, 0), BLOCK, 0) != hipSuccess ||
        per <= 0)
        return 0;
    // cooperative launch: stay well under the co-residency measured for these
    // kernels (6 per CU); 2 per CU gives the relax twice the waves of 1 per CU
    if (persistent) per = per < 2 ? per : 2;
    return (uint32_t)cus * (uint32_t)per;
}

extern "C" int ggak_ms_persistent(gga_dev_graph g, gga_dev_ms ms, uint32_t step0, uint32_t max_steps, int count,
                                  uint32_t grid, uint32_t spin_limit_ms, hipStream_t s) {
    if (grid == 0) return (int)hipErrorInvalidValue;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;
    uint64_t spin_ticks = (uint64_t)spin_limit_ms * (uint64_t)khz;
    void *args[] = {&g, &ms, &step0, &max_steps, &spin_ticks};
    hipError_t e = hipLaunchCooperativeKernel(pick(1, count, ms.lane_major), dim3(grid), dim3(BLOCK), args, 0, s);
    if (e != hipSuccess) return (int)e;
    return launch_status();
}

extern "C" int ggak_ms_harvest(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, gga_ms_outputs out,
                               hipStream_t s) {
    if (lanes == 0) return 0;
    const uint32_t grid = (uint32_t)(((uint64_t)g.n + HV - 1) / HV);
    if (ms.lane_major) k_ms_harvest_lm<<<dim3(grid_for(g.n, 256), LANES), BLOCK, 0, s>>>(g, ms, lanes, out);
    else k_ms_harvest<<<grid, BLOCK, 0, s>>>(g, ms, lanes, out);
    return launch_status();
}

extern "C" int ggak_ms_reach(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, hipStream_t s) {
    // REACH partial slots are always all written (idle workgroups store 0)
    if (ms.lane_major) k_ms_reach_lm<<<dim3(REACH / LANES, LANES), BLOCK, 0, s>>>(g, ms, lanes);
    else k_ms_reach<<<REACH, BLOCK, 0, s>>>(g, ms, lanes);
    return launch_status();
}

extern "C" int ggak_ms_checksum(gga_dev_graph g, gga_dev_ms ms, uint32_t lanes, uint64_t *sums, hipStream_t s) {
    if (lanes == 0) return 0;
    if (ms.lane_major) k_ms_checksum_lm<<<dim3(grid_for(g.n, 64), LANES), BLOCK, 0, s>>>(g, ms, lanes, sums);
    else k_ms_checksum<<<grid_for((uint64_t)g.n * LANES, 1024), BLOCK, 0, s>>>(g, ms, lanes, sums);
    return launch_status();
}

extern "C" int ggak_ms_fold(gga_dev_ms ms, hipStream_t s) {
    k_ms_fold<<<1, BLOCK, 0, s>>>(ms);
    return launch_status();
}
