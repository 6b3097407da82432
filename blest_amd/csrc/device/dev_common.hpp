// Device-side building blocks shared by the traversal kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blest {
namespace dev {

constexpr int kWave = 64;

// Direction and running totals of a level-synchronous traversal.  Level L's expand
// kernel reads slot (L-1)&1 and its block 0 writes slot L&1, so readers and the
// writer never share a slot within one kernel.
//
// A kernel's level argument counts the enqueued levels; the BFS distance it works on is
// that plus `shift`, the levels the fused small-level kernels ran between enqueued levels
// (single-source BFS only, 0 elsewhere).  Slots, counters and frontier buffers follow the
// enqueued level, so the shift travels inside the state the kernels already read.
struct LevelState {
    unsigned dir;        // 0 top-down, 1 bottom-up (the level that wrote the slot)
    unsigned prevNf;     // frontier size that level consumed
    unsigned visited;    // vertices (or vertex-root pairs) reached before that level
    unsigned td, bu;     // levels run in each direction
    unsigned shift;      // distance of the level minus its enqueued level
    unsigned long long cumEdges;  // degree sum of the vertices reached before that level
    unsigned long long buMask;    // bit l-1: level l ran bottom-up (l <= 64)
};

// Control block of one traversal (device memory).  Counters are rotated over three
// slots indexed by level % 3: level L reads slot L%3 (what level L-1 produced),
// writes slot (L+1)%3 and clears slot (L+2)%3, which nobody touches during L.
struct alignas(256) Ctrl {
    unsigned done;                // the traversal has finished (set by an expand kernel)
    unsigned levels;              // number of levels (max distance + 1)
    unsigned tdLevels, buLevels;
    unsigned visited;
    unsigned iterations;
    unsigned levelBase;           // first level of the running replayed chunk (graph mode)
    unsigned histTop;             // multi-source: last level whose frontier words are complete
    unsigned fusedLevels;         // levels minus the enqueued levels the traversal used
    unsigned tailFrom;            // single-source: first enqueued level of the final run of
                                  // small top-down levels with a shrinking frontier (0: none)
    unsigned listCount[3];        // frontier list length (vsets) per slot
    unsigned newCount[3];         // vertices (or vertex-root pairs) discovered per slot
    unsigned long long frontEdges[3];  // degree sum of the discovered vertices per slot
    unsigned long long visitedEdges;
    unsigned long long buMask;
    LevelState state[2];
};

// Streamed once per level: load the slice stream non-temporally so that it does not
// evict the randomly accessed per-vertex state from L2 / the Infinity Cache.
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 stream_load_quad(const uint4* rows, uint32_t q) {
    const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(rows) + q);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t stream_load_u32(const uint32_t* p, uint32_t i) {
    return __builtin_nontemporal_load(p + i);
}

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
// excl[0] == 0 (owner search of the load-balanced expansion).  All lanes must call.
__device__ __forceinline__ unsigned wave_owner(unsigned excl, unsigned t) {
    unsigned lo = 0;
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1) {
        unsigned probe = __shfl(excl, lo + step, kWave);
        if (probe <= t) lo += step;
    }
    return lo;
}

__device__ __forceinline__ unsigned ld_relaxed(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool test_bit(const unsigned long long* bm, unsigned v) {
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

// Block-level aggregation of per-wave counts, flushed with one global atomic per
// counter per block at the end of the kernel.
struct BlockCounts {
    unsigned n;
    unsigned long long e;
};

__device__ __forceinline__ void block_counts_init(BlockCounts* s) {
    if (threadIdx.x == 0) {
        s->n = 0;
        s->e = 0;
    }
    __syncthreads();
}

__device__ __forceinline__ void block_counts_add(BlockCounts* s, unsigned n, unsigned long long e) {
    if (lane_id() == 0 && n) {
        atomicAdd(&s->n, n);
        atomicAdd(&s->e, e);
    }
}

__device__ __forceinline__ void block_counts_flush(BlockCounts* s, unsigned* n, unsigned long long* e) {
    __syncthreads();
    if (threadIdx.x == 0 && s->n) {
        atomicAdd(n, s->n);
        atomicAdd(e, s->e);
    }
}

}  // namespace dev
}  // namespace blest
