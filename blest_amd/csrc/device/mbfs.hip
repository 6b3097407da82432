// Multi-source BFS: up to 128 independent traversals advanced together (gfx950,
// wave64).  The reference declares this plugin (MBFS/MBFSKernel.cuh:21-53) but its
// kernel file is empty; the contract is "one level array per root".
//
// Every vertex carries NW 64-bit words (NW = 1: 64 roots, NW = 2: 128 roots): bit j of
// visited[v] / frontier[v] says whether root j has reached / just reached v.  Level L is
// two stream-ordered kernels, like the single-source traversal:
//   expand  top-down: a slice (row i, mask m) of rset r contributes OR_{b in m} F[8r+b]
//             to row i: the boolean product of the slice with the 8 x 64*NW frontier
//             block of its rset.  Roots not yet at i are OR-ed into next[i].
//           bottom-up: vertex per thread: the roots that have not reached v pull the
//             OR of their in-neighbours' frontier words, stopping once all are found.
//   update  top-down: vertex per thread: new = next[v]; visited |= new; levels of the
//             new (root, vertex) pairs; the vsets of rsets holding a frontier vertex of
//             any root form the next list.  bottom-up: the list from the new frontier.
// Levels are written root-major: levels[j * n + v].
#include "dgraph.hpp"
#include "level_common.hpp"

namespace blest {
namespace dev {

struct MbfsArgs {
    int32_t* levels;
    unsigned long long* visited;
    unsigned long long* front0;
    unsigned long long* front1;
    unsigned long long* next;
    uint32_t* list0;
    uint32_t* list1;
    const uint32_t* srcs;
    uint32_t k;
    Ctrl* ctrl;
    int direction;
    double alpha, beta;
};

template <int NW>
struct Bits {
    unsigned long long w[NW];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = 0;
    }
    __device__ __forceinline__ bool any() const {
        unsigned long long x = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) x |= w[i];
        return x != 0;
    }
    __device__ __forceinline__ unsigned popc() const {
        unsigned c = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) c += __popcll(w[i]);
        return c;
    }
    __device__ __forceinline__ void load(const unsigned long long* p, uint32_t v) {
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = p[size_t(v) * NW + i];
    }
    __device__ __forceinline__ void store(unsigned long long* p, uint32_t v) const {
#pragma unroll
        for (int i = 0; i < NW; ++i) p[size_t(v) * NW + i] = w[i];
    }
};

// Wave-cooperative: lane 0 carries the wave's 64-vertex activity word `bits` (vertices
// [64 word, 64 word + 64)); the vsets of every rset with an active vertex are appended.
__device__ __forceinline__ void mbfs_enqueue(const GraphView& g, uint32_t word, unsigned long long bits,
                                             uint32_t* list, unsigned* counter) {
    bits = __shfl(bits, 0, kWave);
    if (!bits) return;
    const unsigned l = lane_id();
    unsigned mine = 0;
    if (l < 8 && ((bits >> (8 * l)) & 0xFFull)) mine = g.rsetUnit[word * 8 + l + 1] - g.rsetUnit[word * 8 + l];
    const unsigned incl = wave_inclusive_scan(mine);
    const unsigned total = __shfl(incl, kWave - 1, kWave);
    if (total == 0) return;
    unsigned base = 0;
    if (l == 0) base = atomicAdd(counter, total);
    base = __shfl(base, 0, kWave);
    if (mine) {
        const uint32_t r = word * 8 + l;
        const unsigned off = base + incl - mine;
        for (uint32_t u = 0; u < mine; ++u) list[off + u] = g.rsetUnit[r] + u;
    }
}

template <int NW>
__device__ __forceinline__ Bits<NW> root_mask(uint32_t k) {
    Bits<NW> m;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int cnt = int(k) - 64 * i;
        m.w[i] = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    }
    return m;
}

template <int NW>
__global__ void __launch_bounds__(256) mbfs_init_kernel(GraphView g, MbfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    for (size_t i = tid; i < size_t(g.n) * NW; i += nthreads) {
        a.visited[i] = 0ull;
        a.front1[i] = 0ull;
        a.next[i] = 0ull;
    }
    if (tid == 0) level_state_reset(a.ctrl);
}

// Roots set their bits and level 0; counters of the first frontier.
template <int NW>
__global__ void mbfs_seed_kernel(GraphView g, MbfsArgs a) {
    const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.k) return;
    const uint32_t v = a.srcs[j];
    const unsigned long long bit = 1ull << (j & 63);
    atomicOr(&a.visited[size_t(v) * NW + j / 64], bit);
    atomicOr(&a.front1[size_t(v) * NW + j / 64], bit);
    a.levels[size_t(j) * g.n + v] = 0;
    atomicAdd(&a.ctrl->newCount[1], 1u);
    if (g.csrPtr) atomicAdd(&a.ctrl->frontEdges[1], (unsigned long long)(g.csrPtr[v + 1] - g.csrPtr[v]));
}

// The vset list of a frontier given as per-vertex words (first level; after bottom-up).
template <int NW>
__device__ __forceinline__ void mbfs_build_list(const GraphView& g, const unsigned long long* f, uint32_t* list,
                                                unsigned* counter) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t v0 = wave * kWave; v0 < g.W * 64; v0 += nwaves * kWave) {
        const uint32_t v = v0 + lane_id();
        Bits<NW> b;
        b.zero();
        if (v < g.n) b.load(f, v);
        mbfs_enqueue(g, v0 / 64, __ballot(b.any()), list, counter);
    }
}

template <int NW>
__global__ void __launch_bounds__(256) mbfs_first_list_kernel(GraphView g, MbfsArgs a) {
    mbfs_build_list<NW>(g, a.front1, a.list1, &a.ctrl->listCount[1]);
}

template <int NW>
__global__ void __launch_bounds__(256) mbfs_expand_kernel(GraphView g, MbfsArgs a, unsigned L) {
    __shared__ BlockCounts sc;
    Ctrl* c = a.ctrl;
    DirectionParams dp;
    dp.direction = a.direction;
    dp.canBottomUp = g.csrPtr != nullptr;
    dp.alpha = a.alpha;
    dp.beta = a.beta;
    dp.totalEdges = (unsigned long long)a.k * g.nnz;
    dp.totalItems = double(g.n) * a.k;
    bool bottomUp = false;
    if (!level_begin(c, L, dp, bottomUp)) return;
    const unsigned lane = lane_id();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    const unsigned long long* fin = (L & 1) ? a.front1 : a.front0;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    if (!bottomUp) {
        const uint32_t* lin = (L & 1) ? a.list1 : a.list0;
        const unsigned cnt = ld_relaxed(&c->listCount[L % 3]);
        for (unsigned base = wave * kWave; base < cnt; base += nwaves * kWave) {
            const unsigned e = base + lane;
            uint32_t q0 = 0, qc = 0, r = 0;
            if (e < cnt) {
                const uint32_t vs = lin[e];
                q0 = g.unitQuad[vs];
                qc = g.unitQuad[vs + 1] - q0;
                r = g.unitRset[vs];
            }
            const unsigned incl = wave_inclusive_scan(qc);
            const unsigned excl = incl - qc;
            const unsigned total = __shfl(incl, kWave - 1, kWave);
            for (unsigned t0 = 0; t0 < total; t0 += kWave) {
                const unsigned t = t0 + lane;
                const unsigned owner = wave_owner(excl, min(t, total - 1));
                const uint32_t oq0 = __shfl(q0, owner, kWave);
                const unsigned oex = __shfl(excl, owner, kWave);
                const uint32_t orr = __shfl(r, owner, kWave);
                if (t < total) {
                    const uint32_t q = oq0 + (t - oex);
                    const uint4 r4 = g.rows[q];
                    const uint32_t m = g.masks[q];
                    const uint32_t um = (m | (m >> 8) | (m >> 16) | (m >> 24)) & 0xFF;
                    Bits<NW> F[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        F[b].zero();
                        if ((um >> b) & 1) F[b].load(fin, orr * 8 + b);
                    }
                    const uint32_t rowv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const uint32_t ms = (m >> (8 * s)) & 0xFF;
                        if (!ms) continue;
                        Bits<NW> acc;
                        acc.zero();
#pragma unroll
                        for (int b = 0; b < 8; ++b)
                            if ((ms >> b) & 1)
#pragma unroll
                                for (int i = 0; i < NW; ++i) acc.w[i] |= F[b].w[i];
                        if (!acc.any()) continue;
                        const uint32_t row = rowv[s];
#pragma unroll
                        for (int i = 0; i < NW; ++i) {
                            const unsigned long long nv = acc.w[i] & ~a.visited[size_t(row) * NW + i];
                            if (nv) atomicOr(&a.next[size_t(row) * NW + i], nv);
                        }
                    }
                }
            }
        }
        return;
    }
    // bottom-up: vertex per thread
    block_counts_init(&sc);
    const Bits<NW> rootMask = root_mask<NW>(a.k);
    for (uint32_t v0 = wave * kWave; v0 < g.W * 64; v0 += nwaves * kWave) {
        const uint32_t v = v0 + lane;
        Bits<NW> nw;
        nw.zero();
        unsigned long long edges = 0;
        if (v < g.n) {
            Bits<NW> vis, U;
            vis.load(a.visited, v);
            bool need = false;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                U.w[i] = ~vis.w[i] & rootMask.w[i];
                need |= U.w[i] != 0;
            }
            if (need) {
                Bits<NW> acc;
                acc.zero();
                const uint32_t p0 = g.csrPtr[v], p1 = g.csrPtr[v + 1];
                for (uint32_t p = p0; p < p1; ++p) {
                    const uint32_t u = g.csrIdx[p];
                    bool all = true;
#pragma unroll
                    for (int i = 0; i < NW; ++i) {
                        acc.w[i] |= fin[size_t(u) * NW + i];
                        all &= (acc.w[i] & U.w[i]) == U.w[i];
                    }
                    if (all) break;
                }
#pragma unroll
                for (int i = 0; i < NW; ++i) nw.w[i] = acc.w[i] & U.w[i];
                if (nw.any()) {
#pragma unroll
                    for (int i = 0; i < NW; ++i) vis.w[i] |= nw.w[i];
                    vis.store(a.visited, v);
                    for (int i = 0; i < NW; ++i)
                        for (unsigned long long b = nw.w[i]; b; b &= b - 1)
                            a.levels[size_t(64 * i + __builtin_ctzll(b)) * g.n + v] = int32_t(L);
                    edges = (unsigned long long)(p1 - p0) * nw.popc();
                }
            }
            nw.store(fout, v);
        }
        block_counts_add(&sc, wave_sum_u32(nw.popc()), wave_sum_u64(edges));
    }
    const unsigned sOut = (L + 1) % 3;
    block_counts_flush(&sc, &c->newCount[sOut], &c->frontEdges[sOut]);
}

template <int NW>
__global__ void __launch_bounds__(256) mbfs_update_kernel(GraphView g, MbfsArgs a, unsigned L) {
    __shared__ BlockCounts sc;
    Ctrl* c = a.ctrl;
    if (ld_relaxed(&c->done)) return;
    const bool bottomUp = c->state[L & 1].dir != 0;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    uint32_t* lout = (L & 1) ? a.list0 : a.list1;
    const unsigned sOut = (L + 1) % 3;
    if (bottomUp) {
        mbfs_build_list<NW>(g, fout, lout, &c->listCount[sOut]);
        return;
    }
    block_counts_init(&sc);
    const unsigned lane = lane_id();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t v0 = wave * kWave; v0 < g.W * 64; v0 += nwaves * kWave) {
        const uint32_t v = v0 + lane;
        Bits<NW> nw;
        nw.zero();
        if (v < g.n) {
            nw.load(a.next, v);
            if (nw.any()) {
                Bits<NW> vis, z;
                vis.load(a.visited, v);
                z.zero();
                z.store(a.next, v);
#pragma unroll
                for (int i = 0; i < NW; ++i) vis.w[i] |= nw.w[i];
                vis.store(a.visited, v);
            }
            nw.store(fout, v);
        }
        const unsigned cntNew = nw.popc();
        unsigned long long edges = 0;
        if (cntNew) {
            for (int i = 0; i < NW; ++i)
                for (unsigned long long b = nw.w[i]; b; b &= b - 1)
                    a.levels[size_t(64 * i + __builtin_ctzll(b)) * g.n + v] = int32_t(L);
            if (g.csrPtr) edges = (unsigned long long)(g.csrPtr[v + 1] - g.csrPtr[v]) * cntNew;
        }
        mbfs_enqueue(g, v0 / 64, __ballot(cntNew != 0), lout, &c->listCount[sOut]);
        block_counts_add(&sc, wave_sum_u32(cntNew), wave_sum_u64(edges));
    }
    block_counts_flush(&sc, &c->newCount[sOut], &c->frontEdges[sOut]);
}

// Unreached (root, vertex) pairs get -1; root-major rows, so stores are coalesced.
template <int NW>
__global__ void __launch_bounds__(256) mbfs_final_kernel(GraphView g, MbfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    for (uint32_t v = tid; v < g.n; v += nthreads) {
        Bits<NW> vis;
        vis.load(a.visited, v);
        for (uint32_t j = 0; j < a.k; ++j)
            if (!((vis.w[j / 64] >> (j & 63)) & 1ull)) a.levels[size_t(j) * g.n + v] = -1;
    }
}

__global__ void unpermute_rows_kernel(int32_t* out, const int32_t* in, const uint32_t* oldToNew, uint32_t n,
                                      uint32_t rows) {
    for (uint32_t r = blockIdx.y; r < rows; r += gridDim.y)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            out[size_t(r) * n + i] = in[size_t(r) * n + oldToNew[i]];
}

}  // namespace dev

static void mbfs_ensure_state(Query& q, uint32_t nw) {
    DeviceGraph& g = *q.g;
    if (q.mbfsWords >= nw) return;
    const size_t words = size_t(g.n) * nw;
    q.mVis = q.alloc<unsigned long long>(words);
    q.mFront0 = q.alloc<unsigned long long>(words);
    q.mFront1 = q.alloc<unsigned long long>(words);
    q.mNext = q.alloc<unsigned long long>(words);
    if (!q.srcs) q.srcs = q.alloc<uint32_t>(128);
    q.mbfsWords = nw;
}

template <int NW>
static void mbfs_enqueue_levels(Query& q, const dev::MbfsArgs& a) {
    DeviceGraph& g = *q.g;
    const GraphView gv = g.view();
    const dim3 grid = level_grid(g), block(256);
    hipLaunchKernelGGL(dev::mbfs_init_kernel<NW>, grid, block, 0, q.stream, gv, a);
    hipLaunchKernelGGL(dev::mbfs_seed_kernel<NW>, dim3(2), dim3(64), 0, q.stream, gv, a);
    hipLaunchKernelGGL(dev::mbfs_first_list_kernel<NW>, grid, block, 0, q.stream, gv, a);
    run_levels(q, [&](unsigned L) {
        hipLaunchKernelGGL(dev::mbfs_expand_kernel<NW>, grid, block, 0, q.stream, gv, a, L);
        hipLaunchKernelGGL(dev::mbfs_update_kernel<NW>, grid, block, 0, q.stream, gv, a, L);
    });
    hipLaunchKernelGGL(dev::mbfs_final_kernel<NW>, grid, block, 0, q.stream, gv, a);
    BLEST_HIP(hipGetLastError());
}

static void mbfs_run(Query& q, const uint32_t* srcs, uint32_t k, int32_t* hostLevels, const blest_mbfs_opts* o,
                     blest_bfs_stats* st) {
    DeviceGraph& g = *q.g;
    if (k == 0 || k > 128) throw Error(BLEST_ERR_INVALID_ARGUMENT, "multi-source batches hold 1..128 roots");
    if (k > q.maxSources) throw Error(BLEST_ERR_INVALID_ARGUMENT, "more roots than the query's max_sources");
    if (o && o->algo != 0) throw Error(BLEST_ERR_UNSUPPORTED, "only the bit-parallel multi-source kernel is built");
    const uint32_t nw = (k > 64 || (o && o->batch > 64)) ? 2 : 1;
    mbfs_ensure_state(q, nw);
    std::vector<uint32_t> mapped(k);
    for (uint32_t j = 0; j < k; ++j) mapped[j] = map_source(g, srcs[j]);
    BLEST_HIP(hipMemcpyAsync(q.srcs, mapped.data(), 4 * size_t(k), hipMemcpyHostToDevice, q.stream));
    dev::MbfsArgs a;
    a.levels = q.levels;
    a.visited = q.mVis;
    a.front0 = q.mFront0;
    a.front1 = q.mFront1;
    a.next = q.mNext;
    a.list0 = q.list0;
    a.list1 = q.list1;
    a.srcs = q.srcs;
    a.k = k;
    a.ctrl = q.ctrl;
    a.direction = g.hasCsr ? g.opts.direction : 1;
    a.alpha = g.opts.alpha;
    a.beta = g.opts.beta;
    BLEST_HIP(hipEventRecord(q.ev0, q.stream));
    if (nw == 1) mbfs_enqueue_levels<1>(q, a);
    else mbfs_enqueue_levels<2>(q, a);
    BLEST_HIP(hipEventRecord(q.ev1, q.stream));
    BLEST_HIP(hipStreamSynchronize(q.stream));
    dev::Ctrl c;
    BLEST_HIP(hipMemcpy(&c, q.ctrl, sizeof(c), hipMemcpyDeviceToHost));
    q.levelHint = c.levels;
    float ms = 0;
    BLEST_HIP(hipEventElapsedTime(&ms, q.ev0, q.ev1));
    if (st) {
        st->kernel_ms = ms;
        st->total_levels = c.levels;
        st->no_visited = c.visited;
        st->td_levels = c.tdLevels;
        st->bu_levels = c.buLevels;
        st->edges_traversed_component = c.visitedEdges / 2;
        st->bottom_up_mask = c.buMask;
    }
    if (hostLevels) {
        const int32_t* src = q.levels;
        if (g.hasPerm) {
            hipLaunchKernelGGL(dev::unpermute_rows_kernel, dim3(256, std::min<uint32_t>(k, 65535)), dim3(256), 0,
                               q.stream, q.levelsOut, q.levels, g.permOldToNew, g.n, k);
            BLEST_HIP(hipGetLastError());
            src = q.levelsOut;
        }
        BLEST_HIP(hipMemcpyAsync(hostLevels, src, sizeof(int32_t) * size_t(g.n) * k, hipMemcpyDeviceToHost,
                                 q.stream));
        BLEST_HIP(hipStreamSynchronize(q.stream));
    }
}

}  // namespace blest

using namespace blest;

extern "C" blest_status blest_mbfs(blest_query* qh, const uint32_t* srcs, uint32_t k, int32_t* levels,
                                   const blest_mbfs_opts* o, blest_bfs_stats* st) {
    return guarded([&] {
        if (!qh || !srcs) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        mbfs_run(q, srcs, k, levels, o, st);
    });
}
