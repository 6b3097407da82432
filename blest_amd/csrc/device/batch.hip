// Many small graphs in one launch: one workgroup traverses one graph from its root,
// with the visited / frontier / next bitmaps held in LDS (n <= 131072 per graph:
// 3 x 16 KiB), so levels are separated by workgroup barriers only and the bottom-up
// frontier tests never leave the CU.  Top-down streams the graph's BVSS quads
// (slice size 8) for every rset with a frontier byte; bottom-up pulls over the CSR.
#include "dgraph.hpp"

namespace blest {
namespace dev {

struct BatchGraph {
    uint32_t n, R;
    uint64_t levelOff;   // into the concatenated level array
    uint64_t rsetOff;    // into rsetQuad (R + 1 entries per graph, graph-relative quads)
    uint64_t quadOff;    // into rows / masks
    uint64_t csrPtrOff;  // into csrPtr (n + 1 entries per graph, graph-relative)
    uint64_t csrIdxOff;  // into csrIdx
};

struct BatchArgs {
    const BatchGraph* graphs;
    const uint32_t* rsetQuad;
    const uint4* rows;
    const uint32_t* masks;
    const uint32_t* csrPtr;
    const uint32_t* csrIdx;
    const uint32_t* roots;
    const uint32_t* order;
    int32_t* levels;
    uint32_t maxWords;
    double alpha;
};

__global__ void __launch_bounds__(512) batch_bfs_kernel(BatchArgs a) {
    extern __shared__ unsigned long long lds[];
    __shared__ unsigned s_new;
    __shared__ unsigned long long s_edges;
    const uint32_t gi = a.order[blockIdx.x];
    const BatchGraph G = a.graphs[gi];
    const uint32_t W = (G.n + 63) / 64;
    unsigned long long* vis = lds;
    unsigned long long* front = lds + a.maxWords;
    unsigned long long* next = lds + 2 * a.maxWords;
    const uint32_t* rq = a.rsetQuad + G.rsetOff;
    const uint4* rows = a.rows + G.quadOff;
    const uint32_t* masks = a.masks + G.quadOff;
    const uint32_t* cp = a.csrPtr + G.csrPtrOff;
    const uint32_t* ci = a.csrIdx + G.csrIdxOff;
    int32_t* levels = a.levels + G.levelOff;
    const unsigned lane = lane_id(), wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
    const uint32_t src = a.roots[gi];
    for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) {
        const unsigned long long x = (w == src / 64) ? (1ull << (src & 63)) : 0ull;
        vis[w] = x;
        front[w] = x;
        next[w] = 0ull;
    }
    if (threadIdx.x == 0) {
        levels[src] = 0;
        s_new = 1;
        s_edges = cp[src + 1] - cp[src];
    }
    __syncthreads();
    const unsigned long long m = cp[G.n];
    unsigned long long seenEdges = 0;
    bool bottomUp = false;
    for (unsigned L = 1;; ++L) {
        const unsigned nf = s_new;
        const unsigned long long mf = s_edges;
        if (nf == 0) break;
        seenEdges += mf;
        if (!bottomUp && double(mf) * a.alpha > double(m - min(seenEdges, m))) bottomUp = true;
        else if (bottomUp && nf * 24ull < G.n) bottomUp = false;
        __syncthreads();
        if (threadIdx.x == 0) {
            s_new = 0;
            s_edges = 0;
        }
        if (!bottomUp) {
            const uint8_t* fb = reinterpret_cast<const uint8_t*>(front);
            for (uint32_t r = wave; r < G.R; r += nwaves) {
                const uint32_t f = fb[r];
                if (!f) continue;
                for (uint32_t q = rq[r] + lane; q < rq[r + 1]; q += kWave) {
                    const uint4 r4 = rows[q];
                    const uint32_t mk = masks[q];
                    const uint32_t rv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        if ((mk >> (8 * s)) & f) {
                            const uint32_t row = rv[s];
                            if (!((vis[row >> 6] >> (row & 63)) & 1ull)) atomicOr(&next[row >> 6], 1ull << (row & 63));
                        }
                }
            }
        } else {
            for (uint32_t v0 = wave * kWave; v0 < W * 64; v0 += nwaves * kWave) {
                const uint32_t v = v0 + lane;
                bool found = false;
                if (v < G.n && !((vis[v >> 6] >> (v & 63)) & 1ull))
                    for (uint32_t p = cp[v]; p < cp[v + 1]; ++p) {
                        const uint32_t u = ci[p];
                        if ((front[u >> 6] >> (u & 63)) & 1ull) {
                            found = true;
                            break;
                        }
                    }
                const unsigned long long b = __ballot(found);
                if (lane == 0 && b) next[v0 >> 6] = b;
            }
        }
        __syncthreads();
        unsigned cnt = 0;
        unsigned long long edges = 0;
        for (uint32_t w = threadIdx.x; w < W; w += blockDim.x) {
            const unsigned long long nw = next[w] & ~vis[w];
            next[w] = 0ull;
            vis[w] |= nw;
            front[w] = nw;
            for (unsigned long long b = nw; b; b &= b - 1) {
                const uint32_t v = w * 64 + __builtin_ctzll(b);
                levels[v] = int32_t(L);
                edges += cp[v + 1] - cp[v];
                ++cnt;
            }
        }
        cnt = wave_sum_u32(cnt);
        edges = wave_sum_u64(edges);
        if (lane == 0 && cnt) {
            atomicAdd(&s_new, cnt);
            atomicAdd(&s_edges, edges);
        }
        __syncthreads();
    }
    for (uint32_t v = threadIdx.x; v < G.n; v += blockDim.x)
        if (!((vis[v >> 6] >> (v & 63)) & 1ull)) levels[v] = -1;
}

}  // namespace dev

struct GraphBatch {
    int device = 0;
    uint32_t count = 0, maxWords = 0;
    uint64_t totalN = 0;
    std::vector<uint32_t> sizes;
    dev::BatchGraph* graphs = nullptr;
    uint32_t *rsetQuad = nullptr, *masks = nullptr, *csrPtr = nullptr, *csrIdx = nullptr, *roots = nullptr;
    uint32_t* order = nullptr;  // workgroup -> graph
    uint4* rows = nullptr;
    int32_t* levels = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void*> allocs;
    template <class T>
    T* upload(const std::vector<T>& v) {
        void* p = nullptr;
        BLEST_HIP(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)));
        allocs.push_back(p);
        if (!v.empty()) BLEST_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return static_cast<T*>(p);
    }
    ~GraphBatch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocs) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace blest

struct blest_graph_batch {
    blest::GraphBatch b;
};

using namespace blest;

extern "C" {

blest_status blest_graph_batch_create(const blest_csc* const* graphs, uint32_t count, int device,
                                      blest_graph_batch** out) {
    return guarded([&] {
        if (!graphs || !out || count == 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "empty batch");
        auto* h = new blest_graph_batch();
        GraphBatch& gb = h->b;
        try {
            gb.device = device;
            BLEST_HIP(hipSetDevice(device));
            std::vector<dev::BatchGraph> info(count);
            // Per-graph BVSS (variable layout) built in parallel, then concatenated.
            std::vector<BVSS> parts(count);
            std::vector<int> bad(count, 0);
            for (uint32_t i = 0; i < count; ++i)
                if (!graphs[i]) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null graph in batch");
            parallel_dynamic(count, hw_threads(), 1, [&](uint64_t i, unsigned) {
                const CSC& g = csc_of(graphs[i]);
                if (g.n == 0 || g.n > 131072) bad[i] = 1;
                else if (!csc_is_symmetric(g, 1)) bad[i] = 2;
                else parts[i] = bvss_build(g, 8, 4, false, 1);
            });
            for (uint32_t i = 0; i < count; ++i) {
                if (bad[i] == 1) throw Error(BLEST_ERR_UNSUPPORTED, "batched graphs need 1 <= n <= 131072");
                if (bad[i] == 2) throw Error(BLEST_ERR_UNSUPPORTED, "batched graphs must be symmetric");
            }
            uint64_t lv = 0, nr = 0, nq = 0, ncp = 0, nci = 0;
            for (uint32_t i = 0; i < count; ++i) {
                const CSC& g = csc_of(graphs[i]);
                const BVSS& b = parts[i];
                dev::BatchGraph& G = info[i];
                G.n = g.n;
                G.R = b.noRealSliceSets;
                G.levelOff = lv;
                G.rsetOff = nr;
                G.quadOff = nq;
                G.csrPtrOff = ncp;
                G.csrIdxOff = nci;
                lv += g.n;
                nr += uint64_t(b.noRealSliceSets) + 1;
                nq += b.masks.size();
                ncp += g.colPtrs.size();
                nci += g.rows.size();
                gb.sizes.push_back(g.n);
                gb.maxWords = std::max(gb.maxWords, (g.n + 63) / 64);
            }
            std::vector<uint32_t> rq(nr), masks(nq), rowWords(nq * 4), cpAll(ncp), ciAll(nci);
            parallel_dynamic(count, hw_threads(), 16, [&](uint64_t i, unsigned) {
                const CSC& g = csc_of(graphs[i]);
                const BVSS& b = parts[i];
                const dev::BatchGraph& G = info[i];
                for (uint32_t r = 0; r <= b.noRealSliceSets; ++r)
                    rq[G.rsetOff + r] = uint32_t(b.sliceSetPtrs[b.realPtrs[r]] / 4);
                std::copy(b.masks.begin(), b.masks.end(), masks.begin() + G.quadOff);
                std::copy(b.rowIds.begin(), b.rowIds.end(), rowWords.begin() + G.quadOff * 4);
                std::copy(g.colPtrs.begin(), g.colPtrs.end(), cpAll.begin() + G.csrPtrOff);
                std::copy(g.rows.begin(), g.rows.end(), ciAll.begin() + G.csrIdxOff);
            });
            parts.clear();
            // Workgroups start roughly in index order: largest graphs first, so the
            // small ones fill in behind them (longest-processing-time order).
            std::vector<uint32_t> order(count);
            for (uint32_t i = 0; i < count; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return info[x].n > info[y].n; });
            gb.order = gb.upload(order);
            gb.count = count;
            gb.totalN = lv;
            gb.graphs = gb.upload(info);
            gb.rsetQuad = gb.upload(rq);
            gb.masks = gb.upload(masks);
            gb.rows = reinterpret_cast<uint4*>(gb.upload(rowWords));
            gb.csrPtr = gb.upload(cpAll);
            gb.csrIdx = gb.upload(ciAll);
            gb.roots = gb.upload(std::vector<uint32_t>(count, 0));
            gb.levels = gb.upload(std::vector<int32_t>(lv, 0));
            BLEST_HIP(hipStreamCreateWithFlags(&gb.stream, hipStreamNonBlocking));
            BLEST_HIP(hipEventCreate(&gb.ev0));
            BLEST_HIP(hipEventCreate(&gb.ev1));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void blest_graph_batch_free(blest_graph_batch* gb) { delete gb; }

blest_status blest_graph_batch_bfs(blest_graph_batch* h, const uint32_t* roots, int32_t* levels, double* kernel_ms) {
    return guarded([&] {
        if (!h || !roots) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        GraphBatch& gb = h->b;
        BLEST_HIP(hipSetDevice(gb.device));
        for (uint32_t i = 0; i < gb.count; ++i)
            if (roots[i] >= gb.sizes[i]) throw Error(BLEST_ERR_INVALID_ARGUMENT, "root out of range");
        if (levels) check_result_range(levels, 4 * size_t(gb.totalN), gb.device, "levels");
        BLEST_HIP(hipMemcpyAsync(gb.roots, roots, 4 * size_t(gb.count), hipMemcpyHostToDevice, gb.stream));
        dev::BatchArgs a;
        a.graphs = gb.graphs;
        a.rsetQuad = gb.rsetQuad;
        a.rows = gb.rows;
        a.masks = gb.masks;
        a.csrPtr = gb.csrPtr;
        a.csrIdx = gb.csrIdx;
        a.roots = gb.roots;
        a.order = gb.order;
        a.levels = gb.levels;
        a.maxWords = gb.maxWords;
        a.alpha = 15.0;
        const size_t lds = size_t(3) * gb.maxWords * sizeof(unsigned long long);
        BLEST_HIP(hipEventRecord(gb.ev0, gb.stream));
        hipLaunchKernelGGL(dev::batch_bfs_kernel, dim3(gb.count), dim3(512), lds, gb.stream, a);
        BLEST_HIP(hipGetLastError());
        BLEST_HIP(hipEventRecord(gb.ev1, gb.stream));
        if (levels)
            BLEST_HIP(hipMemcpyAsync(levels, gb.levels, 4 * gb.totalN, hipMemcpyDefault, gb.stream));
        BLEST_HIP(hipStreamSynchronize(gb.stream));
        float ms = 0;
        BLEST_HIP(hipEventElapsedTime(&ms, gb.ev0, gb.ev1));
        if (kernel_ms) *kernel_ms = ms;
    });
}

}  // extern "C"
