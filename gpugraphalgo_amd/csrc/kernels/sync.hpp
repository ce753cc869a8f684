// sync.hpp -- workgroup scans and the grid barrier of the persistent kernels.
#pragma once
#include "../gga_dev.h"
#include "wave.hpp"

namespace gga {

// Block-wide exclusive scan of one value per thread; returns the block total.
// s_wave needs blockDim.x / 64 entries.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t &excl, uint32_t *s_wave) {
    const uint32_t wv = threadIdx.x / GGA_WAVE, nw = blockDim.x / GGA_WAVE;
    const uint32_t incl = wave_incl_scan(x);
    if (lane_id() == GGA_WAVE - 1) s_wave[wv] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t i = 0; i < nw; i++) {
        const uint32_t t = s_wave[i];
        off += i < wv ? t : 0u;
        tot += t;
    }
    __syncthreads();
    excl = off + incl - x;
    return tot;
}

// Two-level grid barrier for a cooperative launch: workgroup b arrives at the
// counter of group b % GGA_BAR_GROUPS; the last arrival of a group arrives at
// the top counter; the last group leader releases everyone by bumping the
// generation.  Each counter sees at most ~nblocks / 32 + 32 atomics per
// barrier instead of nblocks.  Every wait is bounded by `spin_ticks` of the
// 100 MHz wall clock; on expiry the error bit is set and every workgroup
// leaves the step loop at its next check.
// Agent-scope fences cost an L2 write-back / invalidate on gfx950, so only
// thread 0 fences: its acq_rel arrival publishes the workgroup's writes (the
// workgroup barrier orders them first), it polls with relaxed coherent loads
// (an acquire load per poll would invalidate L2 every iteration), and one
// acquire fence after the release covers the whole workgroup (shared L1/L2).
__device__ __forceinline__ void grid_barrier(uint32_t *ctl, uint32_t nblocks, uint32_t &gen,
                                             uint64_t spin_ticks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t ngroups = nblocks < GGA_BAR_GROUPS ? nblocks : GGA_BAR_GROUPS;
        const uint32_t grp = blockIdx.x % GGA_BAR_GROUPS;
        const uint32_t gsize = (nblocks - grp + GGA_BAR_GROUPS - 1u) / GGA_BAR_GROUPS;
        bool release = false;
        const uint32_t arrived = __hip_atomic_fetch_add(&ctl[GGA_CTL_BARG(grp)], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == gsize) {
            __hip_atomic_store(&ctl[GGA_CTL_BARG(grp)], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t top = __hip_atomic_fetch_add(&ctl[GGA_CTL_BAR], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_AGENT) + 1u;
            release = top == ngroups;
        }
        if (release) {
            __hip_atomic_store(&ctl[GGA_CTL_BAR], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl[GGA_CTL_BARGEN], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            uint32_t polls = 0;
            while (__hip_atomic_load(&ctl[GGA_CTL_BARGEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (wall_clock64() - t0 > spin_ticks) {
                    __hip_atomic_fetch_or(&ctl[GGA_CTL_ERR], GGA_DEVERR_SPIN, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                // once any workgroup has given up (error bit set), this
                // barrier will not be released: leave now instead of
                // waiting out the full spin limit at every later barrier
                if ((++polls & 63u) == 0u &&
                    __hip_atomic_load(&ctl[GGA_CTL_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        gen += 1u;
    }
    __syncthreads();
}

}  // namespace gga
