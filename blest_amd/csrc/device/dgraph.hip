// HBM-resident device graph and per-query state: the C ABI handles
// blest_device_graph / blest_query, and the host-side level-loop driver.
#include <algorithm>
#include <cstring>

#include "dgraph.hpp"
#include "segments.hpp"

namespace blest {

// ----------------------------------------------------------------- host helpers
DeviceGraph::~DeviceGraph() {
    for (void* p : allocations) (void)hipFree(p);
}

Query::~Query() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& lg : graphs)
        if (lg.exec) (void)hipGraphExecDestroy(lg.exec);
    for (hipEvent_t e : traceEvents) (void)hipEventDestroy(e);
    for (void* p : allocations) (void)hipFree(p);
    if (hostDone) (void)hipHostFree(hostDone);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (evChunk) (void)hipEventDestroy(evChunk);
    if (evIn) (void)hipEventDestroy(evIn);
    if (stream) (void)hipStreamDestroy(stream);
}

dim3 level_grid(const DeviceGraph& g) { return dim3(g.segs); }

dev::SegView DeviceGraph::seg_view() const { return dev::SegView{segs, wps, segBase, chunkBase}; }

dev::SegCounters Query::seg_counters() const { return dev::SegCounters{segCount, segNew, segEdges, segChunks}; }

namespace dev {
// Closes a replayed chunk of `count` levels: the next chunk starts after it.
__global__ void advance_levels_kernel(Ctrl* c, unsigned count) {
    if (threadIdx.x == 0) c->levelBase += count;
}
}  // namespace dev

uint64_t launch_key(double alpha, double beta, int direction, uint64_t extra) {
    uint64_t a, b;
    std::memcpy(&a, &alpha, 8);
    std::memcpy(&b, &beta, 8);
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (uint64_t x : {a, b, uint64_t(uint32_t(direction)), extra}) h = (h ^ x) * 0x100000001B3ull + 0x632BE59BD9B4E019ull;
    return h | 1;  // never 0 (0 = not captured)
}

static void capture_levels(Query& q, LevelGraph& lg, const std::function<void(unsigned)>& levelRel, uint64_t key,
                           const std::function<void(unsigned)>& afterRel) {
    if (lg.exec) {
        BLEST_HIP(hipGraphExecDestroy(lg.exec));
        lg.exec = nullptr;
    }
    const unsigned C = kGraphChunkLevels;
    BLEST_HIP(hipStreamBeginCapture(q.stream, hipStreamCaptureModeThreadLocal));
    for (unsigned k = 0; k < C; ++k) levelRel(k);
    if (afterRel) afterRel(C);
    hipLaunchKernelGGL(dev::advance_levels_kernel, dim3(1), dim3(64), 0, q.stream, q.ctrl, C);
    hipGraph_t graph = nullptr;
    const hipError_t launchErr = hipGetLastError();
    const hipError_t endErr = hipStreamEndCapture(q.stream, &graph);
    if (launchErr != hipSuccess || endErr != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        throw Error(BLEST_ERR_DEVICE, std::string("level graph capture failed: ") +
                                          hipGetErrorString(launchErr != hipSuccess ? launchErr : endErr));
    }
    const hipError_t instErr = hipGraphInstantiate(&lg.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (instErr != hipSuccess) {
        lg.exec = nullptr;
        throw Error(BLEST_ERR_DEVICE, std::string("level graph instantiation failed: ") + hipGetErrorString(instErr));
    }
    lg.chunk = C;
    lg.key = key;
}

static bool poll_done(Query& q, const std::function<void()>& tail) {
    // The tail (e.g. marking unreached vertices) is idempotent and harmless before the
    // end: enqueuing it ahead of the poll hides the host round trip when the chunk was
    // the last one.
    if (tail) tail();
    BLEST_HIP(hipMemcpyAsync(q.hostDone, &q.ctrl->done, sizeof(unsigned), hipMemcpyDeviceToHost, q.stream));
    BLEST_HIP(hipEventRecord(q.evChunk, q.stream));
    BLEST_HIP(hipEventSynchronize(q.evChunk));
    return *q.hostDone != 0;
}

// The first chunk covers the previous traversal's depth (levelHint) plus two levels.  A
// traversal that goes deeper (another root) is usually only a little deeper, so the
// chunks after the first start small and double: a miss costs one more poll, not a
// doubled chunk of levels that only find the done flag (on the 4096-square grid, with
// depths of 60-80 levels, doubling launched up to 230 levels for a 75-level traversal).
void run_levels(Query& q, const std::function<void(unsigned)>& level, const std::function<void()>& tail,
                LevelGraph* lg, const std::function<void(unsigned)>& levelRel, uint64_t key,
                const std::function<void(unsigned)>& after, const std::function<void(unsigned)>& afterRel) {
    const unsigned limit = q.g->n + 2;
    const int mode = q.g->opts.use_graphs;
    if (lg && levelRel && !q.noGraphs && (mode == 1 || (mode == 0 && q.levelHint >= kGraphMinLevels))) {
        if (!lg->exec || lg->key != key) capture_levels(q, *lg, levelRel, key, afterRel);
        const unsigned C = lg->chunk;
        const unsigned maxReps = std::max(1u, 4096u / C);
        unsigned launched = 0;
        unsigned reps = std::max(1u, (q.levelHint + 2 + C - 1) / C), more = 1;
        for (;;) {
            for (unsigned i = 0; i < reps && launched <= limit; ++i, launched += C)
                BLEST_HIP(hipGraphLaunch(lg->exec, q.stream));
            if (poll_done(q, tail)) return;
            if (launched > limit) throw Error(BLEST_ERR_DEVICE, "traversal did not finish within n + 2 levels");
            reps = more;
            more = std::min(more * 2, maxReps);
        }
    }
    unsigned L = 1;
    // with `after` (the fused small-level kernel, which finishes a traversal whose last
    // levels are small): levelHint levels, then `after`; a traversal that needs more
    // level kernels than the last costs one more poll.  (One spare level and a second
    // `after` measured slower: an enqueued level that finds the traversal done still
    // costs ~11 us, and misses were rare.)
    unsigned chunk = after ? q.levelHint : std::max(4u, q.levelHint + 2);
    unsigned more = std::max(after ? 1u : 4u, q.levelHint / 8);
    for (;;) {
        for (unsigned i = 0; i < chunk && L <= limit; ++i) level(L++);
        if (after) after(L);
        if (poll_done(q, tail)) break;
        if (L > limit) throw Error(BLEST_ERR_DEVICE, "traversal did not finish within n + 2 levels");
        chunk = more;
        more = std::min(more * 2, 4096u);
    }
}

bool check_result_range(const void* dst, size_t bytes, int device, const char* what) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, dst) != hipSuccess) {
        (void)hipGetLastError();  // unregistered (pageable) host memory
        return false;
    }
    if (attr.type != hipMemoryTypeDevice) return false;
    if (attr.device != device)
        throw Error(BLEST_ERR_INVALID_ARGUMENT, std::string(what) + ": device memory of another device");
    hipDeviceptr_t base = nullptr;
    size_t extent = 0;
    if (hipMemGetAddressRange(&base, &extent, const_cast<void*>(dst)) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(BLEST_ERR_INVALID_ARGUMENT, std::string(what) + ": not inside a device allocation");
    }
    const char* b = static_cast<const char*>(base);
    const char* d = static_cast<const char*>(dst);
    if (d < b || d + bytes > b + extent)
        throw Error(BLEST_ERR_INVALID_ARGUMENT, std::string(what) + ": needs " + std::to_string(bytes) +
                                                    " bytes, the device allocation holds " +
                                                    std::to_string(size_t(b + extent - d)) + " from there");
    return true;
}

uint32_t map_source(const DeviceGraph& g, uint32_t src) {
    if (src >= g.n) throw Error(BLEST_ERR_INVALID_ARGUMENT, "source vertex out of range");
    return g.hasPerm ? g.hostOldToNew[src] : src;
}

}  // namespace blest

using namespace blest;

extern "C" {

void blest_device_opts_default(blest_device_opts* o) {
    if (!o) return;
    o->device = 0;
    o->block_size = 256;
    o->blocks_per_cu = 4;
    o->alpha = 15.0;
    o->beta = 18.0;
    o->direction = 0;
    o->keep_csr = 1;
    o->timeout_ms = 20000;
    o->use_graphs = 0;
    o->td_quads = 2;
    o->bu_group = 8;
    o->fuse_quads = 2048;
    o->bu_probe = 0;
}

}  // extern "C"

// gh: the CSC the BVSS was built from (ghIsCsr = false; its transpose is the bottom-up
// CSR unless the graph is symmetric), or that transpose itself (ghIsCsr = true, as the
// reference's BVSS keeps it, BVSS::getCSR()).
static blest_status create_device_graph(const blest_bvss* bh, const blest_csc* gh, bool ghIsCsr,
                                        const uint32_t* inv_perm, const blest_device_opts* o,
                                        blest_device_graph** out) {
    return guarded([&] {
        if (!bh || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        const BVSS& b = bvss_of(bh);
        if (b.sliceSize != 8 || b.noMasks != 4)
            throw Error(BLEST_ERR_SLICE_SIZE, "No appropriate kernel found meeting the selected slice size and noMasks.");
        if (b.n == 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "empty graph");
        blest_device_opts opts;
        blest_device_opts_default(&opts);
        if (o) opts = *o;
        if (opts.blocks_per_cu == 0 || opts.blocks_per_cu > 8)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 8]");
        if (opts.keep_csr && !gh)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, ghIsCsr ? "keep_csr needs the CSR" : "keep_csr needs the CSC");
        if (opts.use_graphs < 0 || opts.use_graphs > 2) throw Error(BLEST_ERR_INVALID_ARGUMENT, "use_graphs must be 0, 1 or 2");
        if (opts.td_quads == 0) opts.td_quads = 2;
        if (opts.bu_group == 0) opts.bu_group = 8;
        if (opts.td_quads != 1 && opts.td_quads != 2 && opts.td_quads != 4)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "td_quads must be 1, 2 or 4");
        if (opts.bu_group != 4 && opts.bu_group != 8) throw Error(BLEST_ERR_INVALID_ARGUMENT, "bu_group must be 4 or 8");
        if (opts.fuse_quads == 0) opts.fuse_quads = 2048;
        if (opts.fuse_quads > 65536) throw Error(BLEST_ERR_INVALID_ARGUMENT, "fuse_quads must be at most 65536");
        auto* h = new blest_device_graph();
        DeviceGraph& g = h->g;
        try {
            g.opts = opts;
            g.device = opts.device;
            BLEST_HIP(hipSetDevice(g.device));
            hipDeviceProp_t prop;
            BLEST_HIP(hipGetDeviceProperties(&prop, g.device));
            g.cus = prop.multiProcessorCount;
            g.blockSize = dev::kProducerBlock;
            g.blocksPerCu = opts.blocks_per_cu;
            g.n = b.n;
            g.R = b.noRealSliceSets;
            g.V = b.noVirtualSliceSets;
            g.W = (b.n + 63) / 64;
            g.S = b.noSlices;
            if (b.noSlices / 4 > 0xFFFFFFFFull) throw Error(BLEST_ERR_UNSUPPORTED, "more than 2^32 slice quads");
            // Segments: one producer workgroup per contiguous run of bitmap words.
            g.segs = std::min<uint32_t>(std::min<uint32_t>(uint32_t(g.cus) * g.blocksPerCu, dev::kMaxSegs), g.W);
            g.wps = (g.W + g.segs - 1) / g.segs;
            g.segs = (g.W + g.wps - 1) / g.wps;
            g.gridBlocks = g.segs;
            {
                std::vector<uint32_t> sb(size_t(g.segs) + 1);
                for (uint32_t s = 0; s <= g.segs; ++s) {
                    const uint64_t r = std::min<uint64_t>(uint64_t(s) * g.wps * 8, g.R);
                    sb[s] = b.realPtrs[r];
                }
                g.segBase = g.alloc<uint32_t>(sb.size());
                BLEST_HIP(hipMemcpy(g.segBase, sb.data(), 4 * sb.size(), hipMemcpyHostToDevice));
                // chunk slots: segment s may emit the chunks of its own rsets only
                std::vector<uint32_t> cb(size_t(g.segs) + 1, 0);
                uint64_t total = 0;
                for (uint32_t s = 0; s < g.segs; ++s) {
                    cb[s] = uint32_t(total);
                    const uint64_t r0 = std::min<uint64_t>(uint64_t(s) * g.wps * 8, g.R);
                    const uint64_t r1 = std::min<uint64_t>(uint64_t(s + 1) * g.wps * 8, g.R);
                    for (uint64_t r = r0; r < r1; ++r) total += dev::rset_chunks(b.realPtrs[r + 1] - b.realPtrs[r]);
                }
                if (total > 0xFFFFFFFFull) throw Error(BLEST_ERR_UNSUPPORTED, "more than 2^32 list chunks");
                cb[g.segs] = uint32_t(total);
                g.chunks = total;
                g.chunkBase = g.alloc<uint32_t>(cb.size());
                BLEST_HIP(hipMemcpy(g.chunkBase, cb.data(), 4 * cb.size(), hipMemcpyHostToDevice));
            }
            // Device slice stream: the real slices of each vset packed four to a quad.  The
            // reference layout's padding slots (row 0xFFFFFFFF with an empty mask; a
            // full-padding vset has 128 slots whatever its slice count) carry no edge, and
            // a traversal does not depend on the order of the slices inside a vset, so
            // the device streams real slices only.  Byte s of the mask words is the mask
            // of slot s (slot = lane * 4 + position).
            const uint8_t* slotMask = reinterpret_cast<const uint8_t*>(b.masks.data());
            for (size_t v = 0; v <= g.V; ++v)
                if (b.sliceSetPtrs[v] % 4) throw Error(BLEST_ERR_INVALID_ARGUMENT, "vset not a multiple of 4 slices");
            std::vector<uint32_t> uq(size_t(g.V) + 1, 0);
            const unsigned threads = hw_threads();
            parallel_blocks(g.V, threads, [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t v = lo; v < hi; ++v) {
                    uint32_t real = 0;
                    for (uint64_t sl = b.sliceSetPtrs[v]; sl < b.sliceSetPtrs[v + 1]; ++sl) real += slotMask[sl] != 0;
                    uq[v + 1] = (real + 3) / 4;
                }
            });
            for (size_t v = 0; v < g.V; ++v) uq[v + 1] += uq[v];
            g.Q = uq[g.V];
            std::vector<uint32_t> pRows(size_t(g.Q) * 4, 0xFFFFFFFFu), pMasks(g.Q, 0);
            parallel_blocks(g.V, threads, [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t v = lo; v < hi; ++v) {
                    uint64_t o = uint64_t(uq[v]) * 4;
                    for (uint64_t sl = b.sliceSetPtrs[v]; sl < b.sliceSetPtrs[v + 1]; ++sl) {
                        if (!slotMask[sl]) continue;
                        pRows[o] = b.rowIds[sl];
                        pMasks[o / 4] |= uint32_t(slotMask[sl]) << (8 * (o % 4));
                        ++o;
                    }
                }
            });
            g.unitQuad = g.alloc<uint32_t>(uq.size());
            BLEST_HIP(hipMemcpy(g.unitQuad, uq.data(), 4 * uq.size(), hipMemcpyHostToDevice));
            g.unitRset = g.alloc<uint32_t>(g.V);
            BLEST_HIP(hipMemcpy(g.unitRset, b.virtualToReal.data(), 4 * size_t(g.V), hipMemcpyHostToDevice));
            g.rsetUnit = g.alloc<uint32_t>(size_t(g.R) + 1);
            BLEST_HIP(hipMemcpy(g.rsetUnit, b.realPtrs.data(), 4 * (size_t(g.R) + 1), hipMemcpyHostToDevice));
            g.rows = g.alloc<uint4>(g.Q);
            if (g.Q) BLEST_HIP(hipMemcpy(g.rows, pRows.data(), 16 * size_t(g.Q), hipMemcpyHostToDevice));
            g.masks = g.alloc<uint32_t>(g.Q);
            if (g.Q) BLEST_HIP(hipMemcpy(g.masks, pMasks.data(), 4 * size_t(g.Q), hipMemcpyHostToDevice));
            if (gh) {
                const CSC& c = csc_of(gh);
                if (c.n != b.n)
                    throw Error(BLEST_ERR_INVALID_ARGUMENT, ghIsCsr ? "CSR and BVSS disagree on n"
                                                                    : "CSC and BVSS disagree on n");
                g.nnz = c.nnz();
                g.symmetric = csc_is_symmetric(c);  // a graph is symmetric iff its transpose is
                if (opts.keep_csr) {
                    // Bottom-up reads in-neighbours: the CSR of the edges j -> i, i.e. the
                    // transpose of the CSC (equal to it for a symmetric graph).
                    CSC tr;
                    const CSC* csr = &c;
                    if (!ghIsCsr && !g.symmetric) {
                        tr = csc_transpose(c);
                        csr = &tr;
                    }
                    g.csrPtr = g.alloc<uint32_t>(size_t(c.n) + 1);
                    BLEST_HIP(hipMemcpy(g.csrPtr, csr->colPtrs.data(), 4 * (size_t(c.n) + 1), hipMemcpyHostToDevice));
                    const std::vector<uint32_t>& cp = csr->colPtrs;
                    const std::vector<uint32_t>& ci = csr->rows;
                    g.csrIdx = g.alloc<uint32_t>(csr->nnz());
                    if (opts.bu_probe >= 0) {
                        // Each vertex's in-neighbours by degree, highest first (ties by ID):
                        // high-degree vertices are reached first, so a bottom-up scan meets a
                        // frontier vertex sooner; the first entries also go to `probe`, read
                        // coalesced by a sweep's first probe instead of one random index line
                        // per vertex.  Levels do not depend on the scan order, and the
                        // multi-source sweeps read whole lists.
                        std::vector<uint32_t> sorted(ci.begin(), ci.end()), probe(c.n, 0u);
                        parallel_blocks(c.n, hw_threads(), [&](uint64_t v0, uint64_t v1, unsigned) {
                            for (uint64_t v = v0; v < v1; ++v) {
                                uint32_t* b0 = sorted.data() + cp[v];
                                uint32_t* b1 = sorted.data() + cp[v + 1];
                                std::sort(b0, b1, [&](uint32_t x, uint32_t y) {
                                    const uint32_t dx = cp[x + 1] - cp[x], dy = cp[y + 1] - cp[y];
                                    return dx != dy ? dx > dy : x < y;
                                });
                                probe[v] = b1 > b0 ? *b0 : 0u;
                            }
                        });
                        if (csr->nnz())
                            BLEST_HIP(hipMemcpy(g.csrIdx, sorted.data(), 4 * csr->nnz(), hipMemcpyHostToDevice));
                        g.probe = g.alloc<uint32_t>(c.n);
                        BLEST_HIP(hipMemcpy(g.probe, probe.data(), 4 * size_t(c.n), hipMemcpyHostToDevice));
                    } else if (csr->nnz()) {
                        BLEST_HIP(hipMemcpy(g.csrIdx, ci.data(), 4 * csr->nnz(), hipMemcpyHostToDevice));
                    }
                    g.hasCsr = true;
                    // vertices without in-neighbours: a bottom-up sweep skips them
                    std::vector<unsigned long long> noIn(g.W, 0ull);
                    for (uint32_t v = 0; v < c.n; ++v)
                        if (cp[v + 1] == cp[v]) noIn[v >> 6] |= 1ull << (v & 63);
                    g.noIn = g.alloc<unsigned long long>(g.W);
                    BLEST_HIP(hipMemcpy(g.noIn, noIn.data(), 8 * size_t(g.W), hipMemcpyHostToDevice));
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

extern "C" {

blest_status blest_device_graph_create(const blest_bvss* bh, const blest_csc* gh, const uint32_t* inv_perm,
                                       const blest_device_opts* o, blest_device_graph** out) {
    return create_device_graph(bh, gh, false, inv_perm, o, out);
}

blest_status blest_device_graph_create_csr(const blest_bvss* bh, const blest_csc* csr, const uint32_t* inv_perm,
                                           const blest_device_opts* o, blest_device_graph** out) {
    return create_device_graph(bh, csr, true, inv_perm, o, out);
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
        i->device = g.device;
        i->stream_quads = g.Q;
    });
}

blest_status blest_device_graph_set_thresholds(blest_device_graph* h, double alpha, double beta, int direction) {
    return guarded([&] {
        if (!h) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        if (!(alpha > 0) || !(beta > 0)) throw Error(BLEST_ERR_INVALID_ARGUMENT, "alpha and beta must be positive");
        if (direction < 0 || direction > 2) throw Error(BLEST_ERR_INVALID_ARGUMENT, "direction must be 0, 1 or 2");
        if (direction != 1 && !h->g.hasCsr)
            throw Error(BLEST_ERR_UNSUPPORTED, "bottom-up needs the CSR (keep_csr at creation)");
        h->g.opts.alpha = alpha;
        h->g.opts.beta = beta;
        h->g.opts.direction = direction;
    });
}

blest_status blest_device_graph_set_tiles(blest_device_graph* h, uint32_t td_quads, uint32_t bu_group) {
    return guarded([&] {
        if (!h) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        if (td_quads != 1 && td_quads != 2 && td_quads != 4)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "td_quads must be 1, 2 or 4");
        if (bu_group != 4 && bu_group != 8) throw Error(BLEST_ERR_INVALID_ARGUMENT, "bu_group must be 4 or 8");
        h->g.opts.td_quads = td_quads;
        h->g.opts.bu_group = bu_group;
    });
}

blest_status blest_device_graph_set_fuse(blest_device_graph* h, int32_t fuse_quads) {
    return guarded([&] {
        if (!h) throw Error(BLEST_ERR_INVALID_ARGUMENT, "device graph is null");
        if (fuse_quads > 65536) throw Error(BLEST_ERR_INVALID_ARGUMENT, "fuse_quads must be at most 65536");
        h->g.opts.fuse_quads = fuse_quads ? fuse_quads : 2048;
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
            BLEST_HIP(hipEventCreateWithFlags(&q.evIn, hipEventDisableTiming));
            BLEST_HIP(hipHostMalloc(reinterpret_cast<void**>(&q.hostDone), sizeof(unsigned), hipHostMallocDefault));
            q.ctrl = q.alloc<dev::Ctrl>(1);
            BLEST_HIP(hipMemset(q.ctrl, 0, sizeof(dev::Ctrl)));
            q.levels = q.alloc<int32_t>(size_t(g.n) * max_sources);
            q.visited = q.alloc<unsigned long long>(g.W);
            q.front0 = q.alloc<unsigned long long>(g.W);
            q.front1 = q.alloc<unsigned long long>(g.W);
            q.next = q.alloc<unsigned long long>(g.W);
            q.nextB = q.alloc<uint8_t>(size_t(g.W) * 64);
            q.list0 = q.alloc<uint32_t>(g.V);
            q.list1 = q.alloc<uint32_t>(g.V);
            q.chunk0 = q.alloc<uint2>(g.chunks);
            q.chunk1 = q.alloc<uint2>(g.chunks);
            q.segChunks = q.alloc<uint32_t>(3 * size_t(g.segs));
            q.segCount = q.alloc<uint32_t>(3 * size_t(g.segs));
            q.segNew = q.alloc<uint32_t>(3 * size_t(g.segs));
            q.segEdges = q.alloc<unsigned long long>(3 * size_t(g.segs));
            if (g.hasPerm) q.levelsOut = q.alloc<int32_t>(size_t(g.n) * max_sources);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void blest_query_free(blest_query* q) { delete q; }

blest_status blest_host_alloc(size_t bytes, void** out) {
    return guarded([&] {
        if (!out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if (hipHostMalloc(out, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            throw Error(BLEST_ERR_OUT_OF_HOST_MEMORY, "pinned host allocation failed");
        }
    });
}

void blest_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

blest_status blest_device_memory_info(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        BLEST_HIP(hipSetDevice(device));
        size_t f = 0, t = 0;
        BLEST_HIP(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

// Mirrors blest_query_create plus the lazily allocated multi-source and CC state.
blest_status blest_query_footprint(const blest_device_graph* dg, uint32_t max_sources, uint64_t* bytes) {
    return guarded([&] {
        if (!dg || !bytes) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        const DeviceGraph& g = dg->g;
        if (max_sources == 0) max_sources = 1;
        uint64_t b = sizeof(dev::Ctrl);
        const uint64_t levels = 4ull * g.n * max_sources;
        b += levels * (g.hasPerm ? 2 : 1);
        b += 4ull * 8ull * g.W + 64ull * g.W + 2ull * 4ull * g.V + 2ull * 8ull * g.chunks + 3ull * 20ull * g.segs;
        if (max_sources > 1) {
            const uint32_t nw = max_sources > 64 ? 2 : 1;
            b += (4ull + mbfs_history_levels(g.n, nw) + 1) * 8ull * g.n * nw + 4ull * 128 + 16ull * 128;
            b += 2ull * 8ull * g.W + g.W;  // frontier-any bitmaps, dirty word marks
        }
        b += 2ull * 4ull * g.n;  // CC labels
        *bytes = b;
    });
}


}  // extern "C"

void blest::query_disable_graphs(blest_query* q) {
    if (q) q->q.noGraphs = true;
}
