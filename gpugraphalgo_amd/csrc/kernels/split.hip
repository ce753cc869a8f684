// split.hip -- node splitting as a GPU prefix-sum pass.
//
// Replaces the host-side Graph::NodeSplit (graph.h:415-522),
// getSizeafterLoadBalance (graph.h:386-414), getMaxdeg/getMaxdegIndex
// (graph.h:338-358) and getMaxAllowedOutDegree (graph.h:359-385).  Instead of
// renumbering vertices and adding pseudo-children, the split is a descriptor
// table over the untouched CSR: vertex v of degree d owns
// nch(v) = max(1, ceil(d / T)) consecutive chunk ids starting at
// cstart[v] = sum_{u<v} nch(u) (which is exactly the reference's new parent id
// v + presum[v]); chunk k of v starts at row[v] + k * ceil(d / nch(v))
// (INT_SPLIT sizing, graph.h:461-476) and the last chunk takes the remainder.
//
// HBM traffic: the scan reads row[] twice (reduce + down-sweep) and writes
// cstart[]; the fill reads row[]/cstart[] once and writes cbeg[]/cown[]:
// about 8n + 4n + 8n + 8C bytes in total.
#include "../gga_kernels.h"
#include "wave.hpp"
#include "sync.hpp"

using namespace gga;

namespace {

constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_ITEMS = 16;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 4096 vertices per workgroup

__device__ __forceinline__ uint32_t nchunks(uint32_t deg, uint32_t mda, uint32_t empty) {
    // (deg + mda - 1) / mda without the uint32 overflow; a degree-0 vertex is
    // one node of the split (empty = 1), or has no chunk (empty = 0)
    return deg == 0 ? empty : (deg - 1u) / mda + 1u;
}


__global__ __launch_bounds__(256) void k_degree_max(const uint32_t *__restrict__ row, uint32_t n,
                                                     uint64_t *key_out) {
    uint64_t best = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        uint32_t d = row[v + 1] - row[v];
        uint64_t k = ((uint64_t)d << 32) | (uint64_t)(0xFFFFFFFFu - v);
        best = k > best ? k : best;
    }
    best = wave_max64(best);
    if (lane_id() == 0 && best) atomic_max_u64(key_out, best);
}

__global__ __launch_bounds__(256) void k_degree_hist(const uint32_t *__restrict__ row, uint32_t n,
                                                      uint32_t maxdeg, uint32_t buckets,
                                                      uint32_t *hist) {
    __shared__ uint32_t s_hist[4096];
    for (uint32_t i = threadIdx.x; i < buckets; i += blockDim.x) s_hist[i] = 0;
    __syncthreads();
    const float fm1 = (float)(int)(maxdeg + 1u);
    const float fb = (float)(int)buckets;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        uint32_t d = row[v + 1] - row[v];
        // graph.h:365, each step rounded to float32 exactly as the host code does
        float q = __fdiv_rn((float)d, fm1);
        int idx = (int)__fmul_rn(q, fb);
        if (idx < 0) idx = 0;
        if ((uint32_t)idx >= buckets) idx = (int)buckets - 1;  // unreachable for deg <= maxdeg
        atomicAdd(&s_hist[idx], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < buckets; i += blockDim.x)
        if (s_hist[i]) atomic_add_u32(&hist[i], s_hist[i]);
}

// Pass 1: per-tile sums of nch(v).
__global__ __launch_bounds__(SCAN_THREADS) void k_split_reduce(const uint32_t *__restrict__ row,
                                                               uint32_t n, uint32_t mda, uint32_t empty,
                                                               uint32_t *tile_sum) {
    __shared__ uint32_t s_wave[SCAN_THREADS / GGA_WAVE];
    const uint64_t tile0 = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    for (uint32_t i = 0; i < SCAN_ITEMS; i++) {
        uint64_t v = tile0 + (uint64_t)i * SCAN_THREADS + threadIdx.x;  // coalesced
        if (v < n) s += nchunks(row[v + 1] - row[v], mda, empty);
    }
    uint32_t excl;
    uint32_t tot = block_excl_scan(s, excl, s_wave);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// Pass 2: exclusive scan of the tile sums by one workgroup (carry across 1024-slices).
__global__ __launch_bounds__(1024) void k_split_scan_top(uint32_t *tile_sum, uint32_t ntiles) {
    __shared__ uint32_t s_wave[1024 / GGA_WAVE];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t x = i < ntiles ? tile_sum[i] : 0;
        uint32_t excl;
        uint32_t tot = block_excl_scan(x, excl, s_wave);
        if (i < ntiles) tile_sum[i] = carry + excl;
        carry += tot;
    }
    if (threadIdx.x == 0) tile_sum[ntiles] = carry;
}

// Pass 3: per-thread contiguous 16-vertex runs, block scan, write cstart.
__global__ __launch_bounds__(SCAN_THREADS) void k_split_scan_down(const uint32_t *__restrict__ row,
                                                                  uint32_t n, uint32_t mda, uint32_t empty,
                                                                  const uint32_t *__restrict__ tile_off,
                                                                  uint32_t ntiles,
                                                                  uint32_t *__restrict__ cstart) {
    __shared__ uint32_t s_wave[SCAN_THREADS / GGA_WAVE];
    const uint64_t v0 = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t c[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < SCAN_ITEMS; i++) {
        uint64_t v = v0 + i;
        c[i] = v < n ? nchunks(row[v + 1] - row[v], mda, empty) : 0u;
        s += c[i];
    }
    uint32_t excl;
    block_excl_scan(s, excl, s_wave);
    uint32_t off = tile_off[blockIdx.x] + excl;
#pragma unroll
    for (uint32_t i = 0; i < SCAN_ITEMS; i++) {
        uint64_t v = v0 + i;
        if (v < n) cstart[v] = off;
        off += c[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cstart[n] = tile_off[ntiles];
}

// Chunk descriptors: each wave covers 64 consecutive vertices and writes their
// chunks (a contiguous id range) with all 64 lanes, so a hub's thousands of
// chunks cost its wave a few coalesced rounds rather than one lane a loop.
__global__ __launch_bounds__(256) void k_split_fill(const uint32_t *__restrict__ row,
                                                    const uint32_t *__restrict__ cstart, uint32_t n,
                                                    uint32_t m, uint32_t *__restrict__ cbeg,
                                                    uint32_t *__restrict__ cown, uint32_t *maxchunk) {
    const uint32_t l = lane_id();
    const uint64_t vbase = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) - l;
    const uint64_t v = vbase + l;
    uint32_t cs = 0, nch = 0, rb = 0, local = 0, deg = 0;
    if (v < n) {
        cs = cstart[v];
        nch = cstart[v + 1] - cs;
        rb = row[v];
        deg = row[v + 1] - rb;
        local = deg ? (deg - 1u) / nch + 1u : 0u;  // ceil(deg / nch)
    }
    uint32_t incl = wave_incl_scan(nch);
    uint32_t total = wave_last(incl);
    uint32_t first = wave_bcast(cs, 0);
    for (uint32_t base = 0; base < total; base += GGA_WAVE) {
        uint32_t k = base + l;
        uint32_t kk = k < total ? k : total - 1;
        uint32_t j = wave_owner(incl, kk);
        uint32_t rank = kk - (wave_bcast(incl, j) - wave_bcast(nch, j));
        uint32_t beg = wave_bcast(rb, j) + rank * wave_bcast(local, j);
        if (k < total) {
            cbeg[first + k] = beg;
            cown[first + k] = (uint32_t)vbase + j;
        }
    }
    // largest chunk of a vertex is its first one; reduce per workgroup and
    // stripe over 64 slots (one hot counter would serialize ~16 ns per atomic)
    __shared__ uint32_t s_max[256 / GGA_WAVE];
    const uint64_t mx = wave_max64(local);
    if (l == 0) s_max[threadIdx.x / GGA_WAVE] = (uint32_t)mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = 0;
        for (uint32_t i = 0; i < blockDim.x / GGA_WAVE; i++) b = s_max[i] > b ? s_max[i] : b;
        if (b) atomicMax(&maxchunk[blockIdx.x & (GGA_MAXCHUNK_SLOTS - 1)], b);
    }
    if (v == (uint64_t)n - 1) cbeg[cs + nch] = m;
}

// ---- exclusive scan of a uint32 array (out must not alias in), same three
// passes as the chunk-count scan
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce_arr(const uint32_t *__restrict__ in,
                                                                  uint32_t len, uint32_t *tile_sum) {
    __shared__ uint32_t s_wave[SCAN_THREADS / GGA_WAVE];
    const uint64_t tile0 = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    for (uint32_t i = 0; i < SCAN_ITEMS; i++) {
        const uint64_t v = tile0 + (uint64_t)i * SCAN_THREADS + threadIdx.x;
        if (v < len) s += in[v];
    }
    uint32_t excl;
    const uint32_t tot = block_excl_scan(s, excl, s_wave);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down_arr(const uint32_t *__restrict__ in,
                                                                uint32_t len,
                                                                const uint32_t *__restrict__ tile_off,
                                                                uint32_t ntiles,
                                                                uint32_t *__restrict__ out) {
    __shared__ uint32_t s_wave[SCAN_THREADS / GGA_WAVE];
    const uint64_t v0 = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t c[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < SCAN_ITEMS; i++) {
        const uint64_t v = v0 + i;
        c[i] = v < len ? in[v] : 0u;
        s += c[i];
    }
    uint32_t excl;
    block_excl_scan(s, excl, s_wave);
    uint32_t off = tile_off[blockIdx.x] + excl;
#pragma unroll
    for (uint32_t i = 0; i < SCAN_ITEMS; i++) {
        const uint64_t v = v0 + i;
        if (v < len) out[v] = off;
        off += c[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[len] = tile_off[ntiles];
}

inline int launch_status() { return (int)hipGetLastError(); }

}  // namespace

extern "C" int ggak_scan_u32(const uint32_t *in, uint32_t len, uint32_t *out, uint32_t *d_tmp, hipStream_t s) {
    const uint32_t ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles) k_scan_reduce_arr<<<ntiles, SCAN_THREADS, 0, s>>>(in, len, d_tmp);
    k_split_scan_top<<<1, 1024, 0, s>>>(d_tmp, ntiles);
    if (ntiles) k_scan_down_arr<<<ntiles, SCAN_THREADS, 0, s>>>(in, len, d_tmp, ntiles, out);
    else {
        hipError_t e = hipMemcpyAsync(out, d_tmp, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return (int)e;
    }
    return launch_status();
}

extern "C" int ggak_degree_max(const uint32_t *row, uint32_t n, uint64_t *d_key, hipStream_t s) {
    if (n == 0) return 0;
    uint32_t grid = (n + 255) / 256;
    if (grid > 4096) grid = 4096;
    k_degree_max<<<grid, 256, 0, s>>>(row, n, d_key);
    return launch_status();
}

extern "C" int ggak_degree_hist(const uint32_t *row, uint32_t n, uint32_t maxdeg, uint32_t buckets,
                                uint32_t *d_hist, hipStream_t s) {
    if (n == 0 || buckets == 0 || buckets > 4096) return (int)hipErrorInvalidValue;
    uint32_t grid = (n + 255) / 256;
    if (grid > 2048) grid = 2048;
    k_degree_hist<<<grid, 256, 0, s>>>(row, n, maxdeg, buckets, d_hist);
    return launch_status();
}

extern "C" uint32_t ggak_split_tmp_words(uint32_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

extern "C" int ggak_split_scan_ex(const uint32_t *row, uint32_t n, uint32_t mda, uint32_t empty,
                                  uint32_t *cstart, uint32_t *d_tmp, hipStream_t s) {
    if (mda == 0 || empty > 1u) return (int)hipErrorInvalidValue;
    const uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles) k_split_reduce<<<ntiles, SCAN_THREADS, 0, s>>>(row, n, mda, empty, d_tmp);
    k_split_scan_top<<<1, 1024, 0, s>>>(d_tmp, ntiles);
    if (ntiles) k_split_scan_down<<<ntiles, SCAN_THREADS, 0, s>>>(row, n, mda, empty, d_tmp, ntiles, cstart);
    else {
        // n == 0: cstart[0] = 0 written by the top kernel's total via a copy
        hipError_t e = hipMemcpyAsync(cstart, d_tmp, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return (int)e;
    }
    return launch_status();
}

extern "C" int ggak_split_scan(const uint32_t *row, uint32_t n, uint32_t mda, uint32_t *cstart,
                               uint32_t *d_tmp, hipStream_t s) {
    return ggak_split_scan_ex(row, n, mda, 1u, cstart, d_tmp, s);
}

extern "C" int ggak_split_fill(const uint32_t *row, const uint32_t *cstart, uint32_t n, uint32_t m,
                               uint32_t *cbeg, uint32_t *cown, uint32_t *d_maxchunk, hipStream_t s) {
    if (n == 0) return 0;
    uint64_t threads = ((uint64_t)n + 63) / 64 * 64;
    uint32_t grid = (uint32_t)((threads + 255) / 256);
    k_split_fill<<<grid, 256, 0, s>>>(row, cstart, n, m, cbeg, cown, d_maxchunk);
    return launch_status();
}
