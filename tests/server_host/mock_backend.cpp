// Host-only stand-in for the device side of the request server, for concurrency
// checks of blest_amd/csrc/server.cpp under ThreadSanitizer (no GPU, no HIP).
//
// It implements the C-ABI calls the server makes (query pool, traversals, footprint,
// graph info) and checks the contract the server must keep with them:
//   * a query is used by one thread at a time and never after it is freed;
//   * the bytes of live queries never exceed the server's HBM budget;
//   * every request's result lands in its own output array.
// Results are recognisable patterns: a BFS of root r fills its row with r, a CC fills
// its labels with the graph's tag; per-root statistics encode the root.  Roots equal to
// kFailRoot make the traversal fail, to exercise the error path.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <thread>

#include "blest.h"
#include "capi_internal.hpp"
#include "mock_backend.hpp"

struct blest_device_graph {
    uint32_t n = 0;
    uint32_t symmetric = 1;
    uint32_t tag = 0;
};

struct blest_query {
    blest_device_graph* g = nullptr;
    uint32_t capacity = 1;
    uint64_t bytes = 0;
    std::atomic<int> users{0};
};

namespace blest {
thread_local std::string tl_last_error;
blest_status fail(blest_status s, const std::string& msg) {
    tl_last_error = msg;
    return s;
}
}  // namespace blest

namespace {
std::mutex gMu;
std::set<blest_query*> gLive;  // guarded by gMu
uint64_t gLiveBytes = 0, gPeakBytes = 0, gBudget = 0;
std::atomic<uint64_t> gViolations{0}, gCreated{0}, gCalls{0};

void violation(const char* what) {
    gViolations.fetch_add(1);
    std::fprintf(stderr, "mock backend: %s\n", what);
}

// Marks q in use for the duration of one call; flags overlap and use after free.
struct Use {
    blest_query* q;
    explicit Use(blest_query* q_) : q(q_) {
        {
            std::lock_guard<std::mutex> lk(gMu);
            if (!gLive.count(q)) violation("query used after it was freed");
        }
        if (q->users.fetch_add(1) != 0) violation("query used by two threads at once");
        gCalls.fetch_add(1);
    }
    ~Use() { q->users.fetch_sub(1); }
};

void pause_us(uint32_t salt) {
    thread_local std::mt19937 rng(std::hash<std::thread::id>()(std::this_thread::get_id()) ^ salt);
    std::this_thread::sleep_for(std::chrono::microseconds(rng() % 150));
}

uint64_t footprint(uint32_t cap) { return 1000 + 100ull * cap; }

std::map<uintptr_t, size_t> gDevice;  // registered "device" buffers: start -> bytes (guarded by gMu)

// A result copy of `bytes` to dst: when dst lies in a registered device buffer, the
// whole copy must fit inside it.
void check_dst(const void* dst, size_t bytes) {
    std::lock_guard<std::mutex> lk(gMu);
    const uintptr_t d = reinterpret_cast<uintptr_t>(dst);
    auto it = gDevice.upper_bound(d);
    if (it == gDevice.begin()) return;
    --it;
    if (d >= it->first + it->second) return;  // not inside a device buffer: host memory
    if (d + bytes > it->first + it->second) violation("result copy overruns its device buffer");
}
}  // namespace

// ------------------------------------------------------------------ driver hooks
blest_device_graph* mock_graph_create(uint32_t n, uint32_t symmetric, uint32_t tag) {
    auto* g = new blest_device_graph();
    g->n = n;
    g->symmetric = symmetric;
    g->tag = tag;
    return g;
}
void mock_graph_free(blest_device_graph* g) { delete g; }
void* mock_device_alloc(size_t bytes) {
    void* p = std::malloc(bytes);
    std::lock_guard<std::mutex> lk(gMu);
    gDevice[reinterpret_cast<uintptr_t>(p)] = bytes;
    return p;
}
void mock_device_free(void* p) {
    {
        std::lock_guard<std::mutex> lk(gMu);
        if (!gDevice.erase(reinterpret_cast<uintptr_t>(p))) violation("unknown device buffer freed");
    }
    std::free(p);
}
uint64_t mock_footprint(uint32_t max_sources) { return footprint(max_sources); }
void mock_set_budget(uint64_t bytes) {
    std::lock_guard<std::mutex> lk(gMu);
    gBudget = bytes;
}
mock_report mock_report_get() {
    std::lock_guard<std::mutex> lk(gMu);
    return mock_report{gViolations.load(), gLive.size(), gPeakBytes, gCreated.load(), gCalls.load()};
}

// ------------------------------------------------------------------ C ABI used by the server
extern "C" {
const char* blest_last_error(void) { return blest::tl_last_error.c_str(); }

blest_status blest_device_graph_info_get(const blest_device_graph* dg, blest_device_graph_info* i) {
    *i = blest_device_graph_info{};
    i->n = dg->n;
    i->symmetric = dg->symmetric;
    i->device = 0;
    return BLEST_OK;
}

blest_status blest_device_memory_info(int, uint64_t* free_bytes, uint64_t* total_bytes) {
    *free_bytes = 1ull << 30;
    *total_bytes = 1ull << 31;
    return BLEST_OK;
}

blest_status blest_query_footprint(const blest_device_graph*, uint32_t max_sources, uint64_t* bytes) {
    *bytes = footprint(max_sources);
    return BLEST_OK;
}

blest_status blest_query_create(blest_device_graph* dg, uint32_t max_sources, blest_query** out) {
    auto* q = new blest_query();
    q->g = dg;
    q->capacity = max_sources;
    q->bytes = footprint(max_sources);
    pause_us(1);
    std::lock_guard<std::mutex> lk(gMu);
    gLive.insert(q);
    gLiveBytes += q->bytes;
    gPeakBytes = std::max(gPeakBytes, gLiveBytes);
    if (gBudget && gLiveBytes > gBudget) violation("live query bytes exceed the server's budget");
    gCreated.fetch_add(1);
    *out = q;
    return BLEST_OK;
}

void blest_query_free(blest_query* q) {
    {
        std::lock_guard<std::mutex> lk(gMu);
        if (!gLive.erase(q)) violation("query freed twice");
        gLiveBytes -= q->bytes;
    }
    if (q->users.load() != 0) violation("query freed while in use");
    delete q;
}

blest_status blest_bfs(blest_query* q, uint32_t src, int32_t* levels, blest_bfs_stats* st) {
    Use use(q);
    pause_us(src);
    if (src == kFailRoot) return blest::fail(BLEST_ERR_DEVICE, "planted failure");
    check_dst(levels, sizeof(int32_t) * q->g->n);
    for (uint32_t i = 0; i < q->g->n; ++i) levels[i] = int32_t(src);
    if (st) {
        *st = blest_bfs_stats{};
        st->total_levels = src % 7 + 1;
        st->no_visited = src;
    }
    return BLEST_OK;
}

blest_status blest_cc(blest_query* q, int, uint32_t* labels, blest_cc_stats* st) {
    Use use(q);
    pause_us(q->g->tag);
    check_dst(labels, sizeof(uint32_t) * q->g->n);
    for (uint32_t i = 0; i < q->g->n; ++i) labels[i] = 0xC0000000u | q->g->tag;
    if (st) {
        *st = blest_cc_stats{};
        st->iterations = 3;
        st->no_components = q->g->tag;
    }
    return BLEST_OK;
}
}  // extern "C"

namespace blest {
blest_status mbfs_rows(blest_query* q, const uint32_t* srcs, uint32_t k, int32_t* const* rows, blest_bfs_stats* st,
                       blest_bfs_stats* perRoot) {
    Use use(q);
    if (k > q->capacity) return fail(BLEST_ERR_INVALID_ARGUMENT, "more roots than the query's capacity");
    pause_us(k);
    for (uint32_t j = 0; j < k; ++j)
        if (srcs[j] == kFailRoot) return fail(BLEST_ERR_DEVICE, "planted failure");
    for (uint32_t j = 0; j < k; ++j) check_dst(rows[j], sizeof(int32_t) * q->g->n);
    for (uint32_t j = 0; j < k; ++j)
        for (uint32_t i = 0; i < q->g->n; ++i) rows[j][i] = int32_t(srcs[j]);
    if (st) *st = blest_bfs_stats{};
    for (uint32_t j = 0; j < k; ++j) {
        perRoot[j] = blest_bfs_stats{};
        perRoot[j].total_levels = srcs[j] % 7 + 1;
        perRoot[j].no_visited = srcs[j];
    }
    return BLEST_OK;
}

void query_disable_graphs(blest_query*) {}
}  // namespace blest
