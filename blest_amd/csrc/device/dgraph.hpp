// Host-side handles of the HBM-resident graph and of per-query device state.
#pragma once

#include <algorithm>

#include <hip/hip_runtime.h>

#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../capi_internal.hpp"
#include "dev_common.hpp"
#include "segments.hpp"

namespace blest {

#define BLEST_HIP(call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::blest::Error(e_ == hipErrorOutOfMemory ? BLEST_ERR_OUT_OF_DEVICE_MEMORY    \
                                                           : BLEST_ERR_DEVICE,                 \
                                 std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Read-only view of the device graph passed by value to kernels.
//
// HBM layout: the BVSS of the reference (sliceSetPtrs, virtualToReal, realPtrs, rowIds,
// masks) is uploaded with slice offsets converted to quad offsets: every vset holds a
// multiple of 4 slices (BVSS.cuh:379-415), so its slices are whole "quads" of 4 rows
// (one 16-byte uint4) and one 32-bit mask word (4 x 8-bit slice masks).  A wave64
// streams 64 quads = 256 slices per coalesced step (1 KiB of rows + 256 B of masks).
// Bitmaps (visited, frontier) are 64-bit words: word w holds vertices [64w, 64w+64),
// byte b of the bitmap is the frontier mask of real slice set b.
struct GraphView {
    uint32_t n, R, V, W;          // vertices, rsets, vsets, 64-bit bitmap words
    const uint32_t* unitQuad;     // V + 1: first quad of each vset
    const uint32_t* unitRset;     // V: owning rset (virtualToReal)
    const uint32_t* rsetUnit;     // R + 1: first vset of each rset (realPtrs)
    const uint4* rows;            // Q quads of row IDs (pad rows 0xFFFFFFFF)
    const uint32_t* masks;        // Q mask words
    const uint32_t* csrPtr;       // n + 1, bottom-up in-neighbour lists (nullable)
    const uint32_t* csrIdx;       // nnz
    const unsigned long long* noIn;  // W words: vertices without in-neighbours (nullable)
    const uint32_t* probe;        // n: each vertex's first in-neighbour (its highest-degree one:
                                  // the lists are sorted by degree); nullable
    unsigned long long nnz;       // directed edges
};

struct DeviceGraph {
    int device = 0;
    int cus = 0;
    uint32_t blockSize = 256, blocksPerCu = 4, gridBlocks = 0;
    uint32_t segs = 1, wps = 1;    // producer segments and bitmap words per segment
    uint32_t* segBase = nullptr;   // segs + 1 first list slots
    uint32_t* chunkBase = nullptr; // segs + 1 first chunk slots (rsets of > kChunkVsets vsets)
    uint64_t chunks = 0;           // chunk slots in all
    dev::SegView seg_view() const;
    blest_device_opts opts{};
    uint32_t n = 0, R = 0, V = 0, W = 0;
    uint64_t S = 0, Q = 0, nnz = 0;
    bool symmetric = false, hasCsr = false, hasPerm = false;
    uint32_t *unitQuad = nullptr, *unitRset = nullptr, *rsetUnit = nullptr, *masks = nullptr;
    uint4* rows = nullptr;
    uint32_t *csrPtr = nullptr, *csrIdx = nullptr;
    unsigned long long* noIn = nullptr;  // bit v: v has no in-neighbour (never found bottom-up)
    uint32_t* probe = nullptr;           // first bottom-up probe per vertex
    uint32_t *permNewToOld = nullptr, *permOldToNew = nullptr;  // device copies
    std::vector<uint32_t> hostOldToNew, hostNewToOld;
    std::vector<uint32_t> hostDegree;  // per (reordered) vertex, for GTEPS accounting
    uint64_t deviceBytes = 0;
    std::vector<void*> allocations;

    GraphView view() const {
        GraphView g;
        g.n = n;
        g.R = R;
        g.V = V;
        g.W = W;
        g.unitQuad = unitQuad;
        g.unitRset = unitRset;
        g.rsetUnit = rsetUnit;
        g.rows = rows;
        g.masks = masks;
        g.csrPtr = csrPtr;
        g.csrIdx = csrIdx;
        g.noIn = noIn;
        g.probe = probe;
        g.nnz = nnz;
        return g;
    }
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        BLEST_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        allocations.push_back(p);
        deviceBytes += count * sizeof(T);
        return static_cast<T*>(p);
    }
    ~DeviceGraph();
};

// A captured chunk of levels (kernels that read their level from ctrl->levelBase),
// replayed until the traversal finishes.  key = the launch parameters it was captured
// with; a different key recaptures.
struct LevelGraph {
    hipGraphExec_t exec = nullptr;
    unsigned chunk = 0;
    uint64_t key = 0;
};
enum LevelGraphKind { kGraphBfs = 0, kGraphMbfs1 = 1, kGraphMbfs2 = 2, kGraphCount = 3 };
constexpr unsigned kGraphChunkLevels = 16;  // levels per replayed graph
constexpr unsigned kGraphMinLevels = 64;    // auto mode: replay once a traversal was this deep

// Mutable per-query device state.
struct Query {
    DeviceGraph* g = nullptr;
    uint32_t maxSources = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evChunk = nullptr;
    hipEvent_t evIn = nullptr;     // a caller's stream position that the next launch waits for
    unsigned* hostDone = nullptr;  // pinned: the traversal's done flag, polled per chunk
    unsigned levelHint = 0;        // levels of the last traversal (sizes the first chunk)
    dev::Ctrl* ctrl = nullptr;
    int32_t* levels = nullptr;  // n * maxSources
    unsigned long long *visited = nullptr, *front0 = nullptr, *front1 = nullptr, *next = nullptr;
    uint8_t* nextB = nullptr;  // single-source top-down discoveries, a byte per vertex (W * 64)
    uint32_t *list0 = nullptr, *list1 = nullptr;
    uint2 *chunk0 = nullptr, *chunk1 = nullptr;        // chunk lists (single-source BFS)
    uint32_t *segCount = nullptr, *segNew = nullptr;  // [3][segs]
    uint32_t* segChunks = nullptr;                     // [3][segs]
    unsigned long long* segEdges = nullptr;            // [3][segs]
    dev::SegCounters seg_counters() const;
    int32_t* levelsOut = nullptr;  // original-order copy (when the graph carries a permutation)
    uint32_t* labels = nullptr;    // CC
    uint32_t* labelsAux = nullptr;
    // multi-source state, allocated on first use: words per vertex x n
    uint32_t mbfsWords = 0;
    unsigned long long *mVis = nullptr, *mFront0 = nullptr, *mFront1 = nullptr, *mNext = nullptr;
    unsigned long long* mHist = nullptr;  // frontier words of levels 0..mHistLevels
    unsigned long long* mFull = nullptr;  // rows every root has reached (1 bit per vertex)
    unsigned long long* mAny = nullptr;   // 2 x W: vertices in the frontier of some root (1 bit), by level parity
    uint8_t* mDirty = nullptr;            // W: words whose `next` words were set in the running level
    uint32_t mHistLevels = 0;
    uint32_t* srcs = nullptr;  // multi-source roots (reordered IDs), 128 entries
    void* rootStats = nullptr;  // per-root totals of a multi-source run, 128 entries
    LevelGraph graphs[kGraphCount];
    bool noGraphs = false;  // never replay captured level chunks (queries of the request server)
    std::vector<hipEvent_t> traceEvents;  // one per traced kernel boundary (created on demand)
    std::vector<void*> allocations;
    std::mutex mu;
    uint32_t lastSource = 0;
    bool pending = false;

    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        BLEST_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        allocations.push_back(p);
        return static_cast<T*>(p);
    }
    void release(void* p) {
        if (!p) return;
        auto it = std::find(allocations.begin(), allocations.end(), p);
        if (it == allocations.end()) return;
        (void)hipStreamSynchronize(stream);
        (void)hipFree(p);
        allocations.erase(it);
    }
    ~Query();
};

// Level-loop driver: enqueues level(L) for L = 1, 2, ... in chunks and polls the
// device's done flag after each chunk.  With lg and levelRel given (levelRel(k) launches
// level levelBase + k), deep traversals replay a captured chunk of levels instead
// (device opts use_graphs: 0 auto, 1 always, 2 never; queries with noGraphs never).
dim3 level_grid(const DeviceGraph& g);
// after / afterRel (nullable): a kernel enqueued after each chunk of levels, given the
// next level (absolute / relative to the replayed chunk); with it the first chunk is the
// previous traversal's levelHint levels instead of levelHint + 2.
void run_levels(Query& q, const std::function<void(unsigned)>& level, const std::function<void()>& tail = nullptr,
                LevelGraph* lg = nullptr, const std::function<void(unsigned)>& levelRel = nullptr,
                uint64_t key = 0, const std::function<void(unsigned)>& after = nullptr,
                const std::function<void(unsigned)>& afterRel = nullptr);
uint64_t launch_key(double alpha, double beta, int direction, uint64_t extra);
uint32_t map_source(const DeviceGraph& g, uint32_t src);
// Levels of multi-source frontier words kept for the final level expansion: up to 16,
// within 32 GiB of history per query.
inline uint32_t mbfs_history_levels(uint32_t n, uint32_t nw) {
    const uint64_t perLevel = uint64_t(n) * nw * 8;
    const uint64_t fit = (uint64_t(32) << 30) / std::max<uint64_t>(perLevel, 1);
    return uint32_t(std::min<uint64_t>(16, std::max<uint64_t>(fit, 2) - 1));
}
// Result destinations.  A caller's array may be host memory or device memory of the
// graph's device; when it is device memory, [dst, dst + bytes) must lie inside one
// allocation (hipMemGetAddressRange), else BLEST_ERR_INVALID_ARGUMENT is thrown before
// anything is copied: an undersized or wrong-graph buffer becomes an error, not an
// out-of-bounds device write.  Host memory cannot be measured and is taken as given.
// Returns whether dst is device memory.
bool check_result_range(const void* dst, size_t bytes, int device, const char* what);
// trace (nullable): record an event after the init kernel and after every level kernel.
void bfs_launch(Query& q, uint32_t srcReordered, bool trace = false);
void bfs_finish(Query& q, int32_t* hostLevels, blest_bfs_stats* st);

}  // namespace blest

struct blest_device_graph {
    blest::DeviceGraph g;
};
struct blest_query {
    blest::Query q;
};
