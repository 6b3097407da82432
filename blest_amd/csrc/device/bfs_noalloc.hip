// Single-source BFS that allocates nothing: the per-traversal state (control block,
// bitmaps, discovery bytes, frontier lists and counters, reordered levels) lives in one
// file-scope device array instead of a blest_query's hipMalloc'd buffers, and the
// traversal runs on the caller's stream.  For callers that cannot allocate device
// memory: code paths that must not call hipMalloc (tight serving loops, harnesses that
// forbid it).  The kernels are those of blest_bfs (bfs_launch); only where their state
// lives differs.
//
// Every call on a GPU shares that GPU's copy of the array and leaves its contents
// behind; each call re-initialises what a traversal reads before writing it (the
// control block here, the bitmaps, discovery bytes and counters in bfs_init_kernel).
// Calls on one GPU are serialised by a per-device lock held until the traversal has
// finished, because overlapping calls would race on the shared array.
#include <array>
#include <mutex>

#include "dgraph.hpp"

namespace blest {
namespace {

// 64 MiB per GPU: enough for the state of a scale-22 R-MAT graph (4.2 M vertices,
// 1.1 M vsets: ~31 MiB), held by each GPU the library runs on until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

// Per-device set-up: the device address of `scratch` (each GPU holds its own copy of
// the array at its own address; the host-side name is only a stand-in) and the events
// the level driver records.
struct DeviceSlot {
    std::once_flag once;
    std::mutex mu;  // serialises calls on this device
    unsigned char* ptr = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evChunk = nullptr;
    hipError_t setupErr = hipSuccess;
    std::string setupWhat;
    unsigned levelHint = 0;  // depth of the last traversal (sizes the level driver's first chunk)
};
std::array<DeviceSlot, kMaxDevices> g_slots;

DeviceSlot& slot_for(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(BLEST_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    DeviceSlot& s = g_slots[size_t(device)];
    std::call_once(s.once, [&] {
        // the device is current (the caller set it)
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            s.setupErr = e != hipSuccess ? e : hipErrorInvalidSymbol;
            s.setupWhat = "hipGetSymbolAddress(scratch)";
            return;
        }
        s.ptr = static_cast<unsigned char*>(p);
        for (hipEvent_t* ev : {&s.ev0, &s.ev1}) {
            e = hipEventCreate(ev);
            if (e != hipSuccess) {
                s.setupErr = e;
                s.setupWhat = "hipEventCreate";
                return;
            }
        }
        e = hipEventCreateWithFlags(&s.evChunk, hipEventDisableTiming);
        if (e != hipSuccess) {
            s.setupErr = e;
            s.setupWhat = "hipEventCreateWithFlags";
        }
    });
    if (s.setupErr != hipSuccess)
        throw Error(BLEST_ERR_DEVICE, s.setupWhat + ": " + hipGetErrorString(s.setupErr));
    return s;
}

// Carve the traversal state out of the scratch array (256-byte aligned pieces).
struct Carver {
    unsigned char* base;
    size_t used = 0;
    template <class T>
    T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
        used += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
        return p;
    }
};

// Lays the query state of blest_query_create (single source, without the multi-source
// and CC parts) over `base`; returns the bytes it needs.  With base == nullptr it only
// measures.
size_t carve_state(const DeviceGraph& g, unsigned char* base, Query& q, bool levelsInScratch) {
    Carver c{base};
    q.ctrl = c.take<dev::Ctrl>(1);
    if (levelsInScratch) q.levels = c.take<int32_t>(g.n);
    q.visited = c.take<unsigned long long>(g.W);
    q.front0 = c.take<unsigned long long>(g.W);
    q.front1 = c.take<unsigned long long>(g.W);
    q.nextB = c.take<uint8_t>(size_t(g.W) * 64);
    q.list0 = c.take<uint32_t>(g.V);
    q.list1 = c.take<uint32_t>(g.V);
    q.chunk0 = c.take<uint2>(g.chunks);
    q.chunk1 = c.take<uint2>(g.chunks);
    q.segChunks = c.take<uint32_t>(3 * size_t(g.segs));
    q.segCount = c.take<uint32_t>(3 * size_t(g.segs));
    q.segNew = c.take<uint32_t>(3 * size_t(g.segs));
    q.segEdges = c.take<unsigned long long>(3 * size_t(g.segs));
    return c.used;
}

// A Query whose buffers, stream and events are borrowed: its destructor must release
// none of them.
struct BorrowedQuery : Query {
    ~BorrowedQuery() {
        stream = nullptr;
        ev0 = ev1 = evChunk = evIn = nullptr;
        hostDone = nullptr;
        allocations.clear();
    }
};

__global__ void clear_ctrl_kernel(dev::Ctrl* c) {
    unsigned* w = reinterpret_cast<unsigned*>(c);
    for (unsigned i = threadIdx.x; i < sizeof(dev::Ctrl) / sizeof(unsigned); i += blockDim.x) w[i] = 0u;
}

__global__ void unpermute_kernel(int32_t* out, const int32_t* in, const uint32_t* oldToNew, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[oldToNew[i]];
}

}  // namespace
}  // namespace blest

using namespace blest;

extern "C" {

blest_status blest_bfs_noalloc(const blest_device_graph* dg, uint32_t src, int32_t* levels, blest_bfs_stats* st,
                               void* stream) {
    return guarded([&] {
        if (!dg || !levels) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        DeviceGraph& g = const_cast<DeviceGraph&>(dg->g);  // read-only for the traversal
        const uint32_t s = map_source(g, src);
        BLEST_HIP(hipSetDevice(g.device));
        hipPointerAttribute_t attr{};
        if (hipPointerGetAttributes(&attr, levels) != hipSuccess || attr.type != hipMemoryTypeDevice ||
            attr.device != g.device) {
            (void)hipGetLastError();
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "levels must be device memory of the graph's device");
        }
        hipDeviceptr_t base = nullptr;
        size_t extent = 0;
        if (hipMemGetAddressRange(&base, &extent, levels) != hipSuccess ||
            reinterpret_cast<const char*>(levels) + sizeof(int32_t) * size_t(g.n) >
                reinterpret_cast<const char*>(base) + extent) {
            (void)hipGetLastError();
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "levels must hold n int32 entries");
        }
        BorrowedQuery q;
        const size_t need = carve_state(g, nullptr, q, g.hasPerm);
        if (need > kScratchBytes)
            throw Error(BLEST_ERR_INVALID_ARGUMENT,
                        "graph needs " + std::to_string(need) + " bytes of traversal state, above the " +
                            std::to_string(kScratchBytes) + "-byte scratch of blest_bfs_noalloc (use blest_bfs)");
        DeviceSlot& slot = slot_for(g.device);
        std::lock_guard<std::mutex> lk(slot.mu);
        unsigned hostDone = 0;
        q.g = &g;
        q.maxSources = 1;
        q.stream = static_cast<hipStream_t>(stream);
        q.ev0 = slot.ev0;
        q.ev1 = slot.ev1;
        q.evChunk = slot.evChunk;
        q.hostDone = &hostDone;
        q.noGraphs = true;  // instantiating a captured graph would allocate
        q.levelHint = slot.levelHint;
        carve_state(g, slot.ptr, q, g.hasPerm);
        if (!g.hasPerm) q.levels = levels;  // reordered IDs are the caller's IDs
        hipLaunchKernelGGL(clear_ctrl_kernel, dim3(1), dim3(64), 0, q.stream, q.ctrl);
        BLEST_HIP(hipGetLastError());
        bfs_launch(q, s);
        if (g.hasPerm) {
            hipLaunchKernelGGL(unpermute_kernel, dim3(1024), dim3(256), 0, q.stream, levels, q.levels, g.permOldToNew,
                               g.n);
            BLEST_HIP(hipGetLastError());
        }
        bfs_finish(q, nullptr, st);
        slot.levelHint = q.levelHint;
    });
}

}  // extern "C"
