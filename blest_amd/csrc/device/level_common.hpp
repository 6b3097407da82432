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
// must call it.  When pre is given it also receives the inclusive prefix of the list
// counts (*total = their sum), and cpre likewise of the chunk counts.  L is the enqueued
// level; *effLevel receives the distance it works on (L + the state's shift), *outState
// the state written for it and *inState the previous level's (every thread gets all).
//
// Everything is read as one batch of independent loads (the done flag, the state of
// level L-1, and each counter row once, thread t taking a contiguous run of segments),
// then reduced and scanned in one LDS exchange: a kernel's preamble is one memory round
// trip, not a chain of them (measured: 8.5 us of a small level's 9-13 us expand kernel went to the
// direction sums and the two prefix loads done one after another).
__device__ __forceinline__ bool level_begin(Ctrl* c, const SegCounters& sc, uint32_t segs, unsigned L,
                                            const DirectionParams& p, bool& bottomUp, unsigned* pre = nullptr,
                                            unsigned* total = nullptr, unsigned* cpre = nullptr,
                                            unsigned* ctotal = nullptr, unsigned* effLevel = nullptr,
                                            LevelState* outState = nullptr, LevelState* inState = nullptr) {
    __shared__ unsigned long long s_tmp[16], s_etmp[16];
    __shared__ unsigned s_tmp32[16], s_ctmp32[16];
    const unsigned t = threadIdx.x, nb = blockDim.x;
    const unsigned per = (segs + nb - 1) / nb;
    const size_t base = size_t(L % 3) * segs;
    const unsigned doneIn = ld_relaxed(&c->done);
    const LevelState in = c->state[(L - 1) & 1];
    const unsigned Le = L + in.shift;
    unsigned long long nf = 0, mf = 0;
    unsigned mine = 0, cmine = 0;  // this thread's running list / chunk counts
    for (unsigned j0 = 0; j0 < per; j0 += 4) {
        uint32_t nv[4], cv[4], kv[4];
        unsigned long long ev[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = min(t * per + j0 + j, segs - 1);  // always in range: loads unconditional
            nv[j] = sc.news[base + i];
            ev[j] = sc.edges[base + i];
            cv[j] = pre ? sc.count[base + i] : 0u;
            kv[j] = cpre ? sc.chunks[base + i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = t * per + j0 + j;
            if (j0 + j < per && i < segs) {
                nf += nv[j];
                mf += ev[j];
                mine += cv[j];
                cmine += kv[j];
                if (pre) pre[i] = mine;
                if (cpre) cpre[i] = cmine;
            }
        }
    }
    // levels enqueued past the end of the traversal stop here, before any reduction
    if (__syncthreads_or(doneIn != 0)) return false;
    // the two sums and the two scans share one exchange through LDS (one barrier, not
    // eight): each wave posts its totals, every thread then adds up the waves before it
    {
        const unsigned lane = lane_id(), wave = t / kWave, nw = nb / kWave;
        const unsigned long long wnf = wave_sum_u64(nf), wmf = wave_sum_u64(mf);
        const unsigned incl = wave_inclusive_scan(mine), cincl = wave_inclusive_scan(cmine);
        if (lane == kWave - 1) {
            s_tmp32[wave] = incl;
            s_ctmp32[wave] = cincl;
        }
        if (lane == 0) {
            s_tmp[wave] = wnf;
            s_etmp[wave] = wmf;
        }
        __syncthreads();
        nf = mf = 0;
        unsigned before = 0, cbefore = 0, tot = 0, ctot = 0;
        for (unsigned i = 0; i < nw; ++i) {
            nf += s_tmp[i];
            mf += s_etmp[i];
            const unsigned x = s_tmp32[i], y = s_ctmp32[i];
            before += i < wave ? x : 0u;
            cbefore += i < wave ? y : 0u;
            tot += x;
            ctot += y;
        }
        if (pre) {
            *total = tot;
            const unsigned ex = before + incl - mine;
            for (unsigned j = 0; j < per && t * per + j < segs; ++j) pre[t * per + j] += ex;
        }
        if (cpre) {
            *ctotal = ctot;
            const unsigned ex = cbefore + cincl - cmine;
            for (unsigned j = 0; j < per && t * per + j < segs; ++j) cpre[t * per + j] += ex;
        }
    }
    // the prefixes are complete and every thread has read state / counters before
    // block 0 writes
    __syncthreads();
    if (nf == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->levels = Le - 1;
            c->fusedLevels = in.shift;
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
    bottomUp = (in.dir & 1u) != 0;
    if (p.direction == 0 && p.canBottomUp) {
        const double mu = double(p.totalEdges - min(cum, p.totalEdges));
        if (!bottomUp) {
            // only while the frontier grows: a shrinking tail frontier whose edges exceed
            // what is left unexplored (edges of unreachable vertices stay unexplored) would
            // otherwise sweep every vertex again
            if (double(mf) * p.alpha > mu && nf > in.prevNf) bottomUp = true;
        } else if (double(nf) * p.beta < p.totalItems && nf < in.prevNf) {
            bottomUp = false;
        }
    } else {
        bottomUp = (p.direction == 2 && p.canBottomUp && Le > 1);
    }
    LevelState out;
    out.dir = bottomUp ? 1u : 0u;
    out.prevNf = unsigned(min(nf, 0xFFFFFFFFull));
    out.visited = unsigned(min(in.visited + nf, 0xFFFFFFFFull));
    out.td = in.td + (bottomUp ? 0u : 1u);
    out.bu = in.bu + (bottomUp ? 1u : 0u);
    out.shift = in.shift;
    out.cumEdges = cum;
    out.buMask = in.buMask | ((bottomUp && Le <= 64) ? (1ull << (Le - 1)) : 0ull);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->state[L & 1] = out;
    if (effLevel) *effLevel = Le;
    if (outState) *outState = out;
    if (inState) *inState = in;
    return true;
}

__device__ __forceinline__ void level_state_reset(Ctrl* c) {
    c->done = 0;
    c->levelBase = 1;
    c->fusedLevels = 0;
    c->tailFrom = 0;
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
