// Device-side building blocks shared by the traversal kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blest {
namespace dev {

constexpr int kWave = 64;

// Control block of one persistent traversal (device memory, zeroed per launch).
// Counters are rotated over three slots indexed by level % 3: the slot a level
// writes is reset two levels earlier, so no extra grid barrier is needed to clear it.
struct alignas(256) Ctrl {
    unsigned barCount;   // grid barrier: arrivals in the current round
    unsigned barGen;     // grid barrier: completed rounds
    unsigned error;      // nonzero: a bounded wait expired; every block leaves
    unsigned done;
    unsigned listCount[3];        // frontier list length (vsets) per slot
    unsigned newCount[3];         // vertices (or vertex-root pairs) discovered per slot
    unsigned long long frontEdges[3];  // sum of degrees of the discovered vertices
    unsigned levels;              // number of levels run
    unsigned tdLevels, buLevels;
    unsigned visited;             // vertices reached (including the source)
    unsigned long long visitedEdges;
    unsigned iterations;
    unsigned pad[13];
};

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned wave_inclusive_scan(unsigned x) {
    const unsigned l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        unsigned y = __shfl_up(x, d, kWave);
        if (l >= unsigned(d)) x += y;
    }
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Largest lane j whose value excl[j] <= t, given excl ascending over lanes with
// excl[0] == 0 (owner search of the load-balanced expansion).
__device__ __forceinline__ unsigned wave_owner(unsigned excl, unsigned t) {
    unsigned lo = 0;
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1) {
        unsigned probe = __shfl(excl, lo + step, kWave);
        if (probe <= t) lo += step;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long now_ticks() { return wall_clock64(); }

__device__ __forceinline__ unsigned ld_acquire(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_relaxed(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for a co-resident grid (launched cooperatively).  Every wait is
// bounded by `timeout` wall-clock ticks: on expiry the block raises ctrl->error and
// leaves; blocks waiting elsewhere see the flag and leave too, so the kernel always
// drains.  `gen` is the number of rounds completed so far (meaningful in thread 0).
// Returns false when the traversal must stop because of an error.
__device__ __forceinline__ bool grid_sync(Ctrl* c, unsigned& gen, unsigned long long timeout) {
    __shared__ unsigned s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned ok = 1;
        __threadfence();
        const unsigned nb = gridDim.x;
        unsigned arrived = atomicAdd(&c->barCount, 1u);
        if (arrived == nb - 1) {
            atomicExch(&c->barCount, 0u);
            __threadfence();
            __hip_atomic_store(&c->barGen, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long deadline = now_ticks() + timeout;
            while (ld_acquire(&c->barGen) == gen) {
                if (ld_relaxed(&c->error)) break;
                if (now_ticks() > deadline) {
                    atomicOr(&c->error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        ++gen;
        __threadfence();
        if (ld_relaxed(&c->error)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

__device__ __forceinline__ bool test_bit(const unsigned long long* bm, unsigned v) {
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

}  // namespace dev
}  // namespace blest
