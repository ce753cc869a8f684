// rates.hip -- gga_rates: the memory rates that bound the relax kernel on
// this GPU, measured with the relax kernel's own access pattern but none of
// its chunk bookkeeping.
//
//   stream   : read 2 x 4 B + write 4 B per item over arrays far larger than
//              the caches (the HBM roofline's achievable share)
//   gather   : per edge, stream col[e] and alt[e], then gather dist[col[e]]
//              (the relax guard) over a 2^scale vertex array; destinations
//              drawn from an R-MAT column distribution (each of `scale` bits
//              set with p = b + d = 0.24, as the generator's), i.e. the
//              skewed in-degrees of the R-MAT workloads
//   relax    : the same plus a device-scope atomicMin where the guard passes
//              (a given share of edges), fire-and-forget as in flag steps
// each unsliced (any workgroup takes any edge) and XCD-sliced (edges grouped
// by an 8-way hash of the destination's 128-byte dist line, workgroup b
// taking only group b % 8, so that each XCD's L2 sees one eighth of dist).
// Prints one JSON object; every rate is the best of 5 timed launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            exit(2);                                                                      \
        }                                                                                 \
    } while (0)

namespace {

__host__ __device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__host__ __device__ inline uint32_t line_part(uint32_t v) { return mix32(v >> 5) >> 29; }

constexpr uint32_t BLOCK = 256;
constexpr uint32_t U = 4;   // edge slots per lane in flight, as the relax kernel

__global__ void k_gen(uint32_t *col, uint32_t *alt, uint32_t m, uint32_t scale, uint32_t pass_pct) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t v = 0;
    for (uint32_t l = 0; l < scale; l++) {
        const uint32_t r = mix32(i * 2654435761u + l * 0x9E3779B9u + 12345u) % 100u;
        v = (v << 1) | (r < 24u ? 1u : 0u);
    }
    col[i] = v;
    alt[i] = (mix32(i ^ 0xABCDEFu) % 100u) < pass_pct ? 1u : 0xFFFFFFF0u;
}

__global__ void k_init(uint32_t *dist, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dist[i] = 1000u + (mix32(i) & 0xFFFFu);
}

__global__ __launch_bounds__(BLOCK) void k_stream(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                                   uint32_t *__restrict__ c, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
        c[i] = __builtin_nontemporal_load(&a[i]) + __builtin_nontemporal_load(&b[i]);
}

// Edges [beg, end) of one range: UNROLL slots per lane, every slot's
// address clamped in bounds before any result is used.
template <bool ATOM>
__device__ __forceinline__ void relax_range(const uint32_t *__restrict__ col, const uint32_t *__restrict__ alt,
                                            uint32_t *dist, uint32_t beg, uint32_t end, uint32_t first,
                                            uint32_t stride, uint32_t *sink) {
    for (uint32_t base = beg + first; base < end; base += stride) {
        uint32_t v[U], a[U], c[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            uint32_t i = base + u * BLOCK + threadIdx.x;
            i = i < end ? i : end - 1;
            v[u] = __builtin_nontemporal_load(&col[i]);
            a[u] = __builtin_nontemporal_load(&alt[i]);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) c[u] = dist[v[u]];
        uint32_t keep = 0;
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const bool ok = base + u * BLOCK + threadIdx.x < end && a[u] < c[u];
            if (ATOM && ok) __hip_atomic_fetch_min(&dist[v[u]], a[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            keep += ok ? c[u] : 0u;
        }
        if (!ATOM && keep == 0xFFFFFFFFu) *sink = keep;   // keeps the gathers live
    }
}

// pbeg[0..8]: the sliced groups; [mix_beg, mix_end): edges any workgroup takes.
template <bool ATOM>
__global__ __launch_bounds__(BLOCK) void k_relax(const uint32_t *__restrict__ col, const uint32_t *__restrict__ alt,
                                                 uint32_t *dist, const uint32_t *pbeg, uint32_t mix_beg,
                                                 uint32_t mix_end, uint32_t *sink, uint32_t *xcc_off) {
    const uint32_t x = blockIdx.x & 7u, nbx = gridDim.x >> 3;
    if (threadIdx.x == 0) {
        uint32_t xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        if ((xcc & 7u) != x) atomicAdd(xcc_off, 1u);
    }
    relax_range<ATOM>(col, alt, dist, pbeg[x], pbeg[x + 1], (blockIdx.x >> 3) * BLOCK * U, nbx * BLOCK * U, sink);
    relax_range<ATOM>(col, alt, dist, mix_beg, mix_end, blockIdx.x * BLOCK * U, gridDim.x * BLOCK * U, sink);
}

}  // namespace

int main(int argc, char **argv) {
    uint32_t scale = 22, pass = 7, medges = 32;
    for (int i = 1; i + 1 < argc; i += 2) {
        if (!strcmp(argv[i], "--scale")) scale = (uint32_t)atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--pass")) pass = (uint32_t)atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--medges")) medges = (uint32_t)atoi(argv[i + 1]);
    }
    if (scale < 10 || scale > 28 || pass > 100 || medges < 1 || medges > 256) {
        fprintf(stderr, "usage: gga_rates [--scale 10..28] [--pass 0..100] [--medges 1..256]\n");
        return 1;
    }
    const uint32_t n = 1u << scale, m = medges << 20;
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    uint32_t *d_col, *d_alt, *d_col2, *d_alt2, *d_dist, *d_pbeg, *d_aux;
    CHECK(hipMalloc(&d_col, 4ull * m));
    CHECK(hipMalloc(&d_alt, 4ull * m));
    CHECK(hipMalloc(&d_col2, 4ull * m));
    CHECK(hipMalloc(&d_alt2, 4ull * m));
    CHECK(hipMalloc(&d_dist, 4ull * n));
    CHECK(hipMalloc(&d_pbeg, 4 * 9));
    CHECK(hipMalloc(&d_aux, 8));
    k_gen<<<(m + 255) / 256, 256>>>(d_col, d_alt, m, scale, pass);
    CHECK(hipGetLastError());
    std::vector<uint32_t> col(m), alt(m);
    CHECK(hipMemcpy(col.data(), d_col, 4ull * m, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(alt.data(), d_alt, 4ull * m, hipMemcpyDeviceToHost));
    // guard passes, and the atomics they must issue at least: one per distinct
    // destination (a second pass on a vertex already lowered fails its guard)
    uint64_t passes = 0, atomics = 0;
    {
        std::vector<uint8_t> hit(n, 0);
        for (uint32_t i = 0; i < m; i++)
            if (alt[i] == 1u) {
                passes++;
                if (!hit[col[i]]) {
                    hit[col[i]] = 1;
                    atomics++;
                }
            }
    }
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    auto best_of = [&](auto &&launch) {
        float best = 1e30f;
        for (int rep = 0; rep < 5; rep++) {
            k_init<<<(n + 255) / 256, 256>>>(d_dist, n);
            CHECK(hipEventRecord(e0));
            launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipGetLastError());
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return best;
    };
    // stream: 2 reads + 1 write of 4 B over the edge arrays (3 x 4 m bytes)
    const uint32_t sgrid = (uint32_t)cus * 8u;
    const float t_stream = best_of([&] { k_stream<<<sgrid, BLOCK>>>(d_col, d_alt, d_col2, m); });
    printf("{\"gpu_cus\": %d, \"scale\": %u, \"edges\": %u, \"guard_pass_pct\": %u, \"guard_passes\": %llu, "
           "\"atomics_at_least\": %llu,\n", cus, scale, m, pass, (unsigned long long)passes,
           (unsigned long long)atomics);
    printf(" \"stream_gbs\": %.1f,\n", 12.0 * m / (t_stream * 1e-3) / 1e9);
    const uint32_t grid = (uint32_t)cus * 8u;   // 2048 workgroups of 256 threads, as the relax kernel
    uint32_t h_off = 0;
    printf(" \"relax\": [");
    bool first = true;
    for (int sliced = 0; sliced < 2; sliced++) {
        std::vector<uint32_t> c2(m), a2(m), pbeg(9, 0);
        const uint32_t msl = sliced ? m : 0u;
        for (uint32_t i = 0; i < msl; i++) pbeg[line_part(col[i]) + 1]++;
        for (int p = 0; p < 8; p++) pbeg[p + 1] += pbeg[p];
        std::vector<uint32_t> pos(pbeg.begin(), pbeg.end() - 1);
        for (uint32_t i = 0; i < msl; i++) {
            const uint32_t p = line_part(col[i]);
            c2[pos[p]] = col[i];
            a2[pos[p]++] = alt[i];
        }
        for (uint32_t i = msl; i < m; i++) {
            c2[i] = col[i];
            a2[i] = alt[i];
        }
        CHECK(hipMemcpy(d_col2, c2.data(), 4ull * m, hipMemcpyHostToDevice));
        CHECK(hipMemcpy(d_alt2, a2.data(), 4ull * m, hipMemcpyHostToDevice));
        CHECK(hipMemcpy(d_pbeg, pbeg.data(), 4 * 9, hipMemcpyHostToDevice));
        for (int atom = 0; atom < 2; atom++) {
            CHECK(hipMemset(d_aux, 0, 8));
            const float t = best_of([&] {
                if (atom) k_relax<true><<<grid, BLOCK>>>(d_col2, d_alt2, d_dist, d_pbeg, msl, m, d_aux, d_aux + 1);
                else k_relax<false><<<grid, BLOCK>>>(d_col2, d_alt2, d_dist, d_pbeg, msl, m, d_aux, d_aux + 1);
            });
            CHECK(hipMemcpy(&h_off, d_aux + 1, 4, hipMemcpyDeviceToHost));
            printf("%s\n  {\"layout\": \"%s\", \"atomics\": %s, \"ms\": %.4f, \"g_edges_per_s\": %.1f, "
                   "\"workgroups_off_their_xcd\": %u}",
                   first ? "" : ",", sliced ? "xcd-sliced" : "unsliced", atom ? "true" : "false", t,
                   m / (t * 1e-3) / 1e9, h_off);
            first = false;
        }
    }
    printf("]}\n");
    CHECK(hipFree(d_col));
    CHECK(hipFree(d_alt));
    CHECK(hipFree(d_col2));
    CHECK(hipFree(d_alt2));
    CHECK(hipFree(d_dist));
    CHECK(hipFree(d_pbeg));
    CHECK(hipFree(d_aux));
    return 0;
}
