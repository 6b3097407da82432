// Single-source direction-optimising BFS over the HBM-resident BVSS (gfx950, wave64),
// plus the device-graph / query handles of the C ABI.
//
// Reference path: BVSSBFSKernel::hostCode and its persistent cooperative kernels
// (BFS/BVSSBFSKernel.cuh:33-284 top-down/bottom-up switching, 531-691 top-down only,
// 1283-1608 host side).  On MI355X a grid-wide barrier costs 7-33 us while a kernel
// boundary on one stream costs ~2.3 us (1.6 us replayed from a graph), so the level
// loop is a stream of small kernels instead of one cooperative launch: two kernels
// per level, every decision taken on the device, the host only polling a pinned
// `done` flag once per chunk of levels.
//
// Level L (frontier F = vertices at distance L-1):
//   expand  every block derives the same direction from the previous level's counts;
//           top-down: each wave takes 64 frontier vsets, scans their quad counts
//             across the wave and streams the union of their quads 64 at a time
//             (every lane busy, loads coalesced); a slice hits when
//             (mask & frontier byte) != 0; rows not yet visited are OR-ed into
//             `next` (64-bit atomics);
//           bottom-up: wave per 64-vertex word, lane per vertex: an unvisited vertex
//             scans its in-neighbours until one is in the frontier; the wave's
//             ballot is the new frontier word (no atomics).
//   update  top-down: thread per bitmap word folds `next` into visited / frontier /
//             levels and appends the vsets of rsets with new bits to the next list;
//           bottom-up: the list is rebuilt from the new frontier bitmap.
// Direction: top-down -> bottom-up when the frontier's edges exceed the unexplored
// edges / alpha; back when the frontier shrinks below n / beta.  Levels are unique
// BFS distances, so every schedule produces the same result.
#include <cstring>

#include "dgraph.hpp"
#include "level_common.hpp"

namespace blest {
namespace dev {

struct BfsArgs {
    int32_t* levels;
    unsigned long long* visited;
    unsigned long long* front0;
    unsigned long long* front1;
    unsigned long long* next;
    uint32_t* list0;
    uint32_t* list1;
    Ctrl* ctrl;
    uint32_t src;
    int direction;
    double alpha, beta;
};

__device__ __forceinline__ uint32_t front_byte(const unsigned long long* f, uint32_t r) {
    return reinterpret_cast<const uint8_t*>(f)[r];
}

// Wave-cooperative append of the vsets of every rset with a nonzero byte in `nw`
// (lane's word w).  All 64 lanes must call it.
__device__ __forceinline__ void enqueue_words(const GraphView& g, uint32_t w, unsigned long long nw,
                                              uint32_t* list, unsigned* counter) {
    unsigned cnt = 0;
    if (nw) {
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if ((nw >> (8 * b)) & 0xFFull) {
                const uint32_t r = w * 8 + b;
                cnt += g.rsetUnit[r + 1] - g.rsetUnit[r];
            }
    }
    const unsigned incl = wave_inclusive_scan(cnt);
    const unsigned total = __shfl(incl, kWave - 1, kWave);
    if (total == 0) return;
    unsigned base = 0;
    if (lane_id() == kWave - 1) base = atomicAdd(counter, total);
    base = __shfl(base, kWave - 1, kWave);
    unsigned pos = base + incl - cnt;
    if (cnt) {
        for (int b = 0; b < 8; ++b)
            if ((nw >> (8 * b)) & 0xFFull) {
                const uint32_t r = w * 8 + b;
                for (uint32_t u = g.rsetUnit[r]; u < g.rsetUnit[r + 1]; ++u) list[pos++] = u;
            }
    }
}

__device__ __forceinline__ void expand_slice(uint32_t row, uint32_t hit, const unsigned long long* visited,
                                             unsigned long long* next) {
    if (hit && !test_bit(visited, row)) atomicOr(&next[row >> 6], 1ull << (row & 63));
}

__global__ void bfs_init_kernel(GraphView g, BfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    const uint32_t sw = a.src >> 6;
    const unsigned long long sb = 1ull << (a.src & 63);
    for (uint32_t w = tid; w < g.W; w += nthreads) {
        const unsigned long long x = (w == sw) ? sb : 0ull;
        a.visited[w] = x;
        a.front1[w] = x;
        a.next[w] = 0ull;
    }
    if (blockIdx.x == 0 && threadIdx.x < kWave) {
        const unsigned lane = lane_id();
        const uint32_t r = a.src >> 3;
        const uint32_t u0 = g.rsetUnit[r], u1 = g.rsetUnit[r + 1];
        for (uint32_t u = u0 + lane; u < u1; u += kWave) a.list1[u - u0] = u;
        if (lane == 0) {
            Ctrl* c = a.ctrl;
            a.levels[a.src] = 0;
            level_state_reset(c);
            c->listCount[1] = u1 - u0;
            c->newCount[1] = 1;
            c->frontEdges[1] = g.csrPtr ? (g.csrPtr[a.src + 1] - g.csrPtr[a.src]) : 0ull;
        }
    }
}

__global__ void __launch_bounds__(256) bfs_expand_kernel(GraphView g, BfsArgs a, unsigned L) {
    __shared__ BlockCounts sc;
    DirectionParams dp;
    dp.direction = a.direction;
    dp.canBottomUp = g.csrPtr != nullptr;
    dp.alpha = a.alpha;
    dp.beta = a.beta;
    dp.totalEdges = g.nnz;
    dp.totalItems = double(g.n);
    Ctrl* c = a.ctrl;
    bool bottomUp = false;
    if (!level_begin(c, L, dp, bottomUp)) return;
    const unsigned lane = lane_id();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned nthreads = gridDim.x * blockDim.x;
    const unsigned wave = tid / kWave, nwaves = nthreads / kWave;
    const unsigned long long* fin = (L & 1) ? a.front1 : a.front0;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    if (!bottomUp) {
        const uint32_t* lin = (L & 1) ? a.list1 : a.list0;
        const unsigned cnt = ld_relaxed(&c->listCount[L % 3]);
        for (unsigned base = wave * kWave; base < cnt; base += nwaves * kWave) {
            const unsigned e = base + lane;
            uint32_t q0 = 0, qc = 0, fb = 0;
            if (e < cnt) {
                const uint32_t v = lin[e];
                q0 = g.unitQuad[v];
                qc = g.unitQuad[v + 1] - q0;
                fb = front_byte(fin, g.unitRset[v]);
                if (!fb) qc = 0;
            }
            const unsigned incl = wave_inclusive_scan(qc);
            const unsigned excl = incl - qc;
            const unsigned total = __shfl(incl, kWave - 1, kWave);
            for (unsigned t0 = 0; t0 < total; t0 += kWave) {
                const unsigned t = t0 + lane;
                const unsigned owner = wave_owner(excl, min(t, total - 1));
                const uint32_t oq0 = __shfl(q0, owner, kWave);
                const unsigned oex = __shfl(excl, owner, kWave);
                const uint32_t of = __shfl(fb, owner, kWave);
                if (t < total) {
                    const uint32_t q = oq0 + (t - oex);
                    const uint4 r4 = g.rows[q];
                    const uint32_t m = g.masks[q];
                    expand_slice(r4.x, m & of, a.visited, a.next);
                    expand_slice(r4.y, (m >> 8) & of, a.visited, a.next);
                    expand_slice(r4.z, (m >> 16) & of, a.visited, a.next);
                    expand_slice(r4.w, (m >> 24) & of, a.visited, a.next);
                }
            }
        }
        return;
    }
    // bottom-up: wave per word, lane per vertex
    block_counts_init(&sc);
    for (uint32_t w = wave; w < g.W; w += nwaves) {
        const unsigned long long vis = a.visited[w];
        const uint32_t u = w * 64 + lane;
        bool found = false;
        uint32_t deg = 0;
        if (u < g.n && !((vis >> lane) & 1ull)) {
            const uint32_t p0 = g.csrPtr[u], p1 = g.csrPtr[u + 1];
            deg = p1 - p0;
            for (uint32_t p = p0; p < p1; ++p)
                if (test_bit(fin, g.csrIdx[p])) {
                    found = true;
                    break;
                }
        }
        const unsigned long long nw = __ballot(found);
        if (lane == 0) {
            if (nw) a.visited[w] = vis | nw;
            fout[w] = nw;
        }
        if (found) a.levels[u] = int32_t(L);
        const unsigned long long edges = wave_sum_u64(found ? deg : 0u);
        block_counts_add(&sc, unsigned(__popcll(nw)), edges);
    }
    const unsigned sOut = (L + 1) % 3;
    block_counts_flush(&sc, &c->newCount[sOut], &c->frontEdges[sOut]);
}

__global__ void __launch_bounds__(256) bfs_update_kernel(GraphView g, BfsArgs a, unsigned L) {
    __shared__ BlockCounts sc;
    Ctrl* c = a.ctrl;
    if (ld_relaxed(&c->done)) return;
    const bool bottomUp = c->state[L & 1].dir != 0;
    const unsigned lane = lane_id();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned nthreads = gridDim.x * blockDim.x;
    const unsigned wave = tid / kWave, nwaves = nthreads / kWave;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    uint32_t* lout = (L & 1) ? a.list0 : a.list1;
    const unsigned sOut = (L + 1) % 3;
    if (bottomUp) {  // the next frontier's vset list, from its bitmap
        for (uint32_t w0 = wave * kWave; w0 < g.W; w0 += nwaves * kWave) {
            const uint32_t w = w0 + lane;
            const unsigned long long nw = (w < g.W) ? fout[w] : 0ull;
            enqueue_words(g, w, nw, lout, &c->listCount[sOut]);
        }
        return;
    }
    block_counts_init(&sc);
    const bool canBU = g.csrPtr != nullptr;
    for (uint32_t w0 = wave * kWave; w0 < g.W; w0 += nwaves * kWave) {
        const uint32_t w = w0 + lane;
        unsigned long long nw = 0;
        if (w < g.W) {
            nw = a.next[w];
            if (nw) {
                a.next[w] = 0ull;
                a.visited[w] |= nw;
            }
            fout[w] = nw;
        }
        unsigned long long edges = 0;
        for (unsigned long long bits = nw; bits; bits &= bits - 1) {
            const uint32_t v = w * 64 + __builtin_ctzll(bits);
            a.levels[v] = int32_t(L);
            if (canBU) edges += g.csrPtr[v + 1] - g.csrPtr[v];
        }
        enqueue_words(g, w, nw, lout, &c->listCount[sOut]);
        block_counts_add(&sc, wave_sum_u32(unsigned(__popcll(nw))), wave_sum_u64(edges));
    }
    block_counts_flush(&sc, &c->newCount[sOut], &c->frontEdges[sOut]);
}

// Unreached vertices get -1 (lane per vertex: coalesced stores).
__global__ void __launch_bounds__(256) bfs_final_kernel(GraphView g, BfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    const unsigned lane = lane_id();
    for (uint32_t w = wave; w < g.W; w += nwaves) {
        const unsigned long long vis = a.visited[w];
        const uint32_t u = w * 64 + lane;
        if (u < g.n && !((vis >> lane) & 1ull)) a.levels[u] = -1;
    }
}

__global__ void unpermute_levels_kernel(int32_t* out, const int32_t* in, const uint32_t* oldToNew, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[oldToNew[i]];
}

}  // namespace dev

// ----------------------------------------------------------------- host helpers
DeviceGraph::~DeviceGraph() {
    for (void* p : allocations) (void)hipFree(p);
}

Query::~Query() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocations) (void)hipFree(p);
    if (hostDone) (void)hipHostFree(hostDone);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (evChunk) (void)hipEventDestroy(evChunk);
    if (stream) (void)hipStreamDestroy(stream);
}

dim3 level_grid(const DeviceGraph& g) { return dim3(uint32_t(g.cus) * g.blocksPerCu); }

void run_levels(Query& q, const std::function<void(unsigned)>& level) {
    unsigned L = 1;
    unsigned chunk = std::max(4u, q.levelHint + 2);
    const unsigned limit = q.g->n + 2;
    for (;;) {
        for (unsigned i = 0; i < chunk && L <= limit; ++i) level(L++);
        BLEST_HIP(hipMemcpyAsync(q.hostDone, &q.ctrl->done, sizeof(unsigned), hipMemcpyDeviceToHost, q.stream));
        BLEST_HIP(hipEventRecord(q.evChunk, q.stream));
        BLEST_HIP(hipEventSynchronize(q.evChunk));
        if (*q.hostDone) break;
        if (L > limit) throw Error(BLEST_ERR_DEVICE, "traversal did not finish within n + 2 levels");
        chunk = std::min(chunk * 2, 4096u);
    }
}

uint32_t map_source(const DeviceGraph& g, uint32_t src) {
    if (src >= g.n) throw Error(BLEST_ERR_INVALID_ARGUMENT, "source vertex out of range");
    return g.hasPerm ? g.hostOldToNew[src] : src;
}

void bfs_launch(Query& q, uint32_t src) {
    DeviceGraph& g = *q.g;
    dev::BfsArgs a;
    a.levels = q.levels;
    a.visited = q.visited;
    a.front0 = q.front0;
    a.front1 = q.front1;
    a.next = q.next;
    a.list0 = q.list0;
    a.list1 = q.list1;
    a.ctrl = q.ctrl;
    a.src = src;
    a.direction = g.hasCsr ? g.opts.direction : 1;
    a.alpha = g.opts.alpha;
    a.beta = g.opts.beta;
    const GraphView gv = g.view();
    const dim3 grid = level_grid(g), block(256);
    BLEST_HIP(hipEventRecord(q.ev0, q.stream));
    hipLaunchKernelGGL(dev::bfs_init_kernel, grid, block, 0, q.stream, gv, a);
    run_levels(q, [&](unsigned L) {
        hipLaunchKernelGGL(dev::bfs_expand_kernel, grid, block, 0, q.stream, gv, a, L);
        hipLaunchKernelGGL(dev::bfs_update_kernel, grid, block, 0, q.stream, gv, a, L);
    });
    hipLaunchKernelGGL(dev::bfs_final_kernel, grid, block, 0, q.stream, gv, a);
    BLEST_HIP(hipGetLastError());
    BLEST_HIP(hipEventRecord(q.ev1, q.stream));
    q.lastSource = src;
    q.pending = true;
}

void bfs_finish(Query& q, int32_t* hostLevels, blest_bfs_stats* st) {
    DeviceGraph& g = *q.g;
    BLEST_HIP(hipStreamSynchronize(q.stream));
    q.pending = false;
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
            hipLaunchKernelGGL(dev::unpermute_levels_kernel, dim3(1024), dim3(256), 0, q.stream, q.levelsOut,
                               q.levels, g.permOldToNew, g.n);
            BLEST_HIP(hipGetLastError());
            src = q.levelsOut;
        }
        BLEST_HIP(hipMemcpyAsync(hostLevels, src, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost, q.stream));
        BLEST_HIP(hipStreamSynchronize(q.stream));
    }
}

}  // namespace blest

using namespace blest;

extern "C" {

void blest_device_opts_default(blest_device_opts* o) {
    if (!o) return;
    o->device = 0;
    o->block_size = 256;
    o->blocks_per_cu = 8;
    o->alpha = 15.0;
    o->beta = 18.0;
    o->direction = 0;
    o->keep_csr = 1;
    o->timeout_ms = 20000;
}

blest_status blest_device_graph_create(const blest_bvss* bh, const blest_csc* gh, const uint32_t* inv_perm,
                                       const blest_device_opts* o, blest_device_graph** out) {
    return guarded([&] {
        if (!bh || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        const BVSS& b = bvss_of(bh);
        if (b.sliceSize != 8 || b.noMasks != 4)
            throw Error(BLEST_ERR_SLICE_SIZE, "No appropriate kernel found meeting the selected slice size and noMasks.");
        if (b.n == 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "empty graph");
        blest_device_opts opts;
        blest_device_opts_default(&opts);
        if (o) opts = *o;
        if (opts.blocks_per_cu == 0 || opts.blocks_per_cu > 64)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 64]");
        if (opts.keep_csr && !gh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "keep_csr needs the CSC");
        auto* h = new blest_device_graph();
        DeviceGraph& g = h->g;
        try {
            g.opts = opts;
            g.device = opts.device;
            BLEST_HIP(hipSetDevice(g.device));
            hipDeviceProp_t prop;
            BLEST_HIP(hipGetDeviceProperties(&prop, g.device));
            g.cus = prop.multiProcessorCount;
            g.blockSize = 256;
            g.blocksPerCu = opts.blocks_per_cu;
            g.gridBlocks = uint32_t(g.cus) * g.blocksPerCu;
            g.n = b.n;
            g.R = b.noRealSliceSets;
            g.V = b.noVirtualSliceSets;
            g.W = (b.n + 63) / 64;
            g.S = b.noSlices;
            g.Q = b.noSlices / 4;
            if (g.Q > 0xFFFFFFFFull) throw Error(BLEST_ERR_UNSUPPORTED, "more than 2^32 slice quads");
            std::vector<uint32_t> uq(size_t(g.V) + 1);
            for (size_t v = 0; v <= g.V; ++v) {
                if (b.sliceSetPtrs[v] % 4) throw Error(BLEST_ERR_INVALID_ARGUMENT, "vset not a multiple of 4 slices");
                uq[v] = uint32_t(b.sliceSetPtrs[v] / 4);
            }
            g.unitQuad = g.alloc<uint32_t>(uq.size());
            BLEST_HIP(hipMemcpy(g.unitQuad, uq.data(), 4 * uq.size(), hipMemcpyHostToDevice));
            g.unitRset = g.alloc<uint32_t>(g.V);
            BLEST_HIP(hipMemcpy(g.unitRset, b.virtualToReal.data(), 4 * size_t(g.V), hipMemcpyHostToDevice));
            g.rsetUnit = g.alloc<uint32_t>(size_t(g.R) + 1);
            BLEST_HIP(hipMemcpy(g.rsetUnit, b.realPtrs.data(), 4 * (size_t(g.R) + 1), hipMemcpyHostToDevice));
            g.rows = g.alloc<uint4>(g.Q);
            BLEST_HIP(hipMemcpy(g.rows, b.rowIds.data(), 16 * g.Q, hipMemcpyHostToDevice));
            g.masks = g.alloc<uint32_t>(g.Q);
            BLEST_HIP(hipMemcpy(g.masks, b.masks.data(), 4 * g.Q, hipMemcpyHostToDevice));
            if (gh) {
                const CSC& c = csc_of(gh);
                if (c.n != b.n) throw Error(BLEST_ERR_INVALID_ARGUMENT, "CSC and BVSS disagree on n");
                g.nnz = c.nnz();
                g.symmetric = csc_is_symmetric(c);
                if (opts.keep_csr) {
                    // Bottom-up reads in-neighbours: the CSR of the edges j -> i, i.e. the
                    // transpose of the CSC (equal to it for a symmetric graph).
                    CSC tr;
                    const CSC* csr = &c;
                    if (!g.symmetric) {
                        tr = csc_transpose(c);
                        csr = &tr;
                    }
                    g.csrPtr = g.alloc<uint32_t>(size_t(c.n) + 1);
                    BLEST_HIP(hipMemcpy(g.csrPtr, csr->colPtrs.data(), 4 * (size_t(c.n) + 1), hipMemcpyHostToDevice));
                    g.csrIdx = g.alloc<uint32_t>(csr->nnz());
                    if (csr->nnz())
                        BLEST_HIP(hipMemcpy(g.csrIdx, csr->rows.data(), 4 * csr->nnz(), hipMemcpyHostToDevice));
                    g.hasCsr = true;
                }
            }
            if (inv_perm) {
                if (!permutation_is_bijection(inv_perm, b.n)) throw Error(BLEST_ERR_PERMUTATION, "Permutation check has failed.");
                g.hostOldToNew.assign(inv_perm, inv_perm + b.n);
                g.hostNewToOld.resize(b.n);
                for (uint32_t old = 0; old < b.n; ++old) g.hostNewToOld[inv_perm[old]] = old;
                g.permOldToNew = g.alloc<uint32_t>(b.n);
                g.permNewToOld = g.alloc<uint32_t>(b.n);
                BLEST_HIP(hipMemcpy(g.permOldToNew, g.hostOldToNew.data(), 4 * size_t(b.n), hipMemcpyHostToDevice));
                BLEST_HIP(hipMemcpy(g.permNewToOld, g.hostNewToOld.data(), 4 * size_t(b.n), hipMemcpyHostToDevice));
                g.hasPerm = true;
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void blest_device_graph_free(blest_device_graph* dg) { delete dg; }

blest_status blest_device_graph_info_get(const blest_device_graph* h, blest_device_graph_info* i) {
    return guarded([&] {
        if (!h || !i) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        const DeviceGraph& g = h->g;
        i->n = g.n;
        i->rsets = g.R;
        i->vsets = g.V;
        i->grid_blocks = g.gridBlocks;
        i->block_size = g.blockSize;
        i->cus = uint32_t(g.cus);
        i->symmetric = g.symmetric;
        i->has_csr = g.hasCsr;
        i->slices = g.S;
        i->nnz = g.nnz;
        i->device_bytes = g.deviceBytes;
    });
}

blest_status blest_query_create(blest_device_graph* dg, uint32_t max_sources, blest_query** out) {
    return guarded([&] {
        if (!dg || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        if (max_sources == 0) max_sources = 1;
        DeviceGraph& g = dg->g;
        BLEST_HIP(hipSetDevice(g.device));
        auto* h = new blest_query();
        Query& q = h->q;
        try {
            q.g = &g;
            q.maxSources = max_sources;
            BLEST_HIP(hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking));
            BLEST_HIP(hipEventCreate(&q.ev0));
            BLEST_HIP(hipEventCreate(&q.ev1));
            BLEST_HIP(hipEventCreateWithFlags(&q.evChunk, hipEventDisableTiming));
            BLEST_HIP(hipHostMalloc(reinterpret_cast<void**>(&q.hostDone), sizeof(unsigned), hipHostMallocDefault));
            q.ctrl = q.alloc<dev::Ctrl>(1);
            BLEST_HIP(hipMemset(q.ctrl, 0, sizeof(dev::Ctrl)));
            q.levels = q.alloc<int32_t>(size_t(g.n) * max_sources);
            q.visited = q.alloc<unsigned long long>(g.W);
            q.front0 = q.alloc<unsigned long long>(g.W);
            q.front1 = q.alloc<unsigned long long>(g.W);
            q.next = q.alloc<unsigned long long>(g.W);
            q.list0 = q.alloc<uint32_t>(g.V);
            q.list1 = q.alloc<uint32_t>(g.V);
            if (g.hasPerm) q.levelsOut = q.alloc<int32_t>(size_t(g.n) * max_sources);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void blest_query_free(blest_query* q) { delete q; }


blest_status blest_bfs(blest_query* qh, uint32_t src, int32_t* levels, blest_bfs_stats* st) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        bfs_launch(q, map_source(*q.g, src));
        bfs_finish(q, levels, st);
    });
}

blest_status blest_bfs_async(blest_query* qh, uint32_t src, void* stream) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        Query& q = qh->q;
        (void)stream;  // the query's own stream orders its launches
        BLEST_HIP(hipSetDevice(q.g->device));
        bfs_launch(q, map_source(*q.g, src));
    });
}

blest_status blest_query_sync(blest_query* qh, blest_bfs_stats* st) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        bfs_finish(qh->q, nullptr, st);
    });
}

void* blest_query_levels_device(blest_query* qh) { return qh ? qh->q.levels : nullptr; }

}  // extern "C"
