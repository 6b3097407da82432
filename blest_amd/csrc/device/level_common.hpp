// Direction decision shared by the level-synchronous traversals (single- and
// multi-source BFS).  Every block of level L's expand kernel evaluates it from the same
// inputs, so all blocks agree; block 0 publishes the result in ctrl->state[L & 1].
#pragma once

#include "dev_common.hpp"
#include "segments.hpp"

namespace blest {
namespace dev {

struct DirectionParams {
    int direction;                  // 0 optimising, 1 top-down only, 2 bottom-up after level 1
    bool canBottomUp;               // the bottom-up CSR is resident
    double alpha, beta;
    unsigned long long totalEdges;  // edges (or edge-root pairs) of the whole graph
    double totalItems;              // vertices (or vertex-root pairs)
};

// Reads what level L-1 produced (per-segment counters of slot L % 3), decides level
// L's direction and returns false when the previous level discovered nothing (then
// block 0 records the final totals and raises ctrl->done).  Every thread of the block
// must call it.
__device__ __forceinline__ bool level_begin(Ctrl* c, const SegCounters& sc, uint32_t segs, unsigned L,
                                            const DirectionParams& p, bool& bottomUp) {
    __shared__ unsigned long long s_tmp[16];
    if (ld_relaxed(&c->done)) return false;
    const LevelState in = c->state[(L - 1) & 1];
    const unsigned sIn = L % 3;
    const unsigned long long nf = block_sum_row_u32(sc.news + size_t(sIn) * segs, segs, s_tmp);
    const unsigned long long mf = block_sum_row_u64(sc.edges + size_t(sIn) * segs, segs, s_tmp);
    __syncthreads();  // every thread has read state / counters before block 0 writes
    if (nf == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->levels = L - 1;
            c->tdLevels = in.td;
            c->buLevels = in.bu;
            c->visited = in.visited;
            c->visitedEdges = in.cumEdges;
            c->buMask = in.buMask;
            __hip_atomic_store(&c->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return false;
    }
    const unsigned long long cum = in.cumEdges + mf;
    bottomUp = in.dir != 0;
    if (p.direction == 0 && p.canBottomUp) {
        const double mu = double(p.totalEdges - min(cum, p.totalEdges));
        if (!bottomUp) {
            if (double(mf) * p.alpha > mu) bottomUp = true;
        } else if (double(nf) * p.beta < p.totalItems && nf < in.prevNf) {
            bottomUp = false;
        }
    } else {
        bottomUp = (p.direction == 2 && p.canBottomUp && L > 1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        LevelState out;
        out.dir = bottomUp ? 1u : 0u;
        out.prevNf = unsigned(min(nf, 0xFFFFFFFFull));
        out.visited = unsigned(min(in.visited + nf, 0xFFFFFFFFull));
        out.td = in.td + (bottomUp ? 0u : 1u);
        out.bu = in.bu + (bottomUp ? 1u : 0u);
        out.pad = 0;
        out.cumEdges = cum;
        out.buMask = in.buMask | ((bottomUp && L <= 64) ? (1ull << (L - 1)) : 0ull);
        c->state[L & 1] = out;
    }
    return true;
}

__device__ __forceinline__ void level_state_reset(Ctrl* c) {
    c->done = 0;
    c->levelBase = 1;
    c->state[0] = LevelState{0, 0, 0, 0, 0, 0, 0ull, 0ull};
    c->state[1] = c->state[0];
}

// Level of a kernel: fixed at launch, or (in a replayed chunk of levels) the chunk's
// base level held in ctrl plus the kernel's position in the chunk.
__device__ __forceinline__ unsigned level_of(const Ctrl* c, int fromCtrl, unsigned L) {
    return fromCtrl ? __hip_atomic_load(&c->levelBase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + L : L;
}

}  // namespace dev
}  // namespace blest
