// Concurrency stress of the request server (blest_amd/csrc/server.cpp) over the
// host-only backend, meant to run under ThreadSanitizer.
//
// Several client threads submit BFS and CC requests on three graphs, each request with
// its own output array, keep a window of requests in flight, and check every result:
// status, batch size, statistics and the array contents.  The server runs more
// workers than clients with an HBM budget of two batch queries plus two single ones,
// so the pool evicts idle queries and workers wait for running ones.  A last phase
// frees the server with requests still queued (the destructor must finish them).
// Exit status 0 and the line "server stress ok" mean every check held.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "blest.h"
#include "mock_backend.hpp"

namespace {

constexpr uint32_t kN = 48;  // vertices per mock graph
std::atomic<int> gFailures{0};

void failf(const char* fmt, unsigned long long a, unsigned long long b) {
    if (gFailures.fetch_add(1) < 20) {
        std::fprintf(stderr, fmt, a, b);
        std::fputc('\n', stderr);
    }
}

struct Pending {
    uint64_t ticket;
    uint32_t graph;
    blest_request_kind kind;
    uint32_t root;
    std::shared_ptr<int32_t> out;  // host memory, or a registered "device" buffer
};

bool gShortBuffers = false;  // planted mistake: device buffers one entry short

std::shared_ptr<int32_t> result_array(uint32_t n, bool device) {
    if (device)
        return std::shared_ptr<int32_t>(
            static_cast<int32_t*>(mock_device_alloc(sizeof(int32_t) * (gShortBuffers ? n - 1 : n))),
                                        [](int32_t* p) { mock_device_free(p); });
    return std::shared_ptr<int32_t>(new int32_t[n], std::default_delete<int32_t[]>());
}

void check_result(blest_server* srv, Pending& p, blest_device_graph_info const* infos) {
    blest_request_result res{};
    const blest_status st = blest_server_wait(srv, p.ticket, 0xFFFFFFFFu, &res);
    const bool planted = p.kind == BLEST_REQ_BFS && p.root == kFailRoot;
    if (planted) {
        if (st != BLEST_ERR_DEVICE) failf("ticket %llu: planted failure gave status %llu", p.ticket, st);
        return;  // a failed batch leaves its rows unspecified
    }
    if (st != BLEST_OK && st != BLEST_ERR_DEVICE) failf("ticket %llu: status %llu", p.ticket, st);
    if (st == BLEST_ERR_DEVICE) return;  // batched with a planted failure
    if (res.ticket != p.ticket) failf("result of ticket %llu names ticket %llu", p.ticket, res.ticket);
    if (res.batch_size == 0 || res.batch_size > 64) failf("ticket %llu: batch size %llu", p.ticket, res.batch_size);
    const uint32_t n = infos[p.graph].n;
    if (p.kind == BLEST_REQ_BFS) {
        if (res.total_levels != p.root % 7 + 1 || res.no_visited != p.root)
            failf("ticket %llu: statistics of another root (levels %llu)", p.ticket, res.total_levels);
        for (uint32_t i = 0; i < n; ++i)
            if (p.out.get()[i] != int32_t(p.root)) {
                failf("ticket %llu: row holds %llu", p.ticket, uint32_t(p.out.get()[i]));
                break;
            }
    } else {
        if (res.total_levels != 3) failf("ticket %llu: CC iterations %llu", p.ticket, res.total_levels);
        for (uint32_t i = 0; i < n; ++i)
            if (uint32_t(p.out.get()[i]) != (0xC0000000u | (p.graph + 1))) {
                failf("ticket %llu: labels hold %llu", p.ticket, uint32_t(p.out.get()[i]));
                break;
            }
    }
}

void client(blest_server* srv, const blest_device_graph_info* infos, uint32_t id, uint32_t requests, bool device) {
    std::mt19937 rng(1234 + id);
    std::deque<Pending> window;
    for (uint32_t i = 0; i < requests; ++i) {
        Pending p;
        p.graph = rng() % 3;
        p.kind = (rng() % 5 == 0 && infos[p.graph].symmetric) ? BLEST_REQ_CC : BLEST_REQ_BFS;
        p.root = (rng() % 401 == 0) ? kFailRoot : rng() % infos[p.graph].n;
        p.out = result_array(infos[p.graph].n, device && (rng() % 4 != 0));  // device phase: 3 in 4
        std::memset(p.out.get(), 0xFF, sizeof(int32_t) * infos[p.graph].n);
        for (;;) {
            const blest_status st = blest_server_submit(srv, p.graph, p.kind, p.root, p.out.get(), &p.ticket);
            if (st == BLEST_OK) break;
            if (st == BLEST_ERR_INVALID_ARGUMENT && p.root == kFailRoot) {
                p.root = 0;  // root out of range for this graph: submit a valid one
                continue;
            }
            if (st != BLEST_ERR_BUSY) {
                failf("client %llu: submit status %llu", id, st);
                return;
            }
            if (!window.empty()) {
                check_result(srv, window.front(), infos);
                window.pop_front();
            } else {
                std::this_thread::yield();
            }
        }
        window.push_back(std::move(p));
        if (window.size() >= 24) {
            check_result(srv, window.front(), infos);
            window.pop_front();
        }
        if (i % 97 == 0 && !window.empty()) {  // poll the newest without collecting it
            blest_request_result res{};
            const blest_status st = blest_server_wait(srv, window.back().ticket, 0, &res);
            if (st == BLEST_OK || st == BLEST_ERR_DEVICE) window.pop_back();  // collected: drop it
            else if (st != BLEST_ERR_TIMEOUT) failf("client %llu: poll status %llu", id, st);
        }
    }
    while (!window.empty()) {
        check_result(srv, window.front(), infos);
        window.pop_front();
    }
}

}  // namespace

int main(int argc, char** argv) {
    const uint32_t clients = 4, requests = argc > 1 ? uint32_t(std::atoi(argv[1])) : 1500;
    gShortBuffers = argc > 2 && std::strcmp(argv[2], "short") == 0;
    // three graphs: two symmetric (BFS + CC), one directed (BFS only); n > kFailRoot on
    // the first so the planted root is in range there
    blest_device_graph* graphs[3] = {mock_graph_create(1024, 1, 1), mock_graph_create(kN, 1, 2),
                                     mock_graph_create(kN + 5, 0, 3)};
    blest_device_graph_info infos[3];
    for (int g = 0; g < 3; ++g) blest_device_graph_info_get(graphs[g], &infos[g]);

    blest_server_opts o;
    blest_server_opts_default(&o);
    o.workers = 6;
    o.max_batch = 16;
    o.min_multi = 3;
    o.capacity = 64;
    o.hbm_budget = 2 * mock_footprint(16) + 2 * mock_footprint(1);
    mock_set_budget(o.hbm_budget);

    blest_server* srv = nullptr;
    if (blest_server_create(graphs, 3, &o, &srv) != BLEST_OK) {
        std::fprintf(stderr, "server create failed: %s\n", blest_last_error());
        return 1;
    }
    std::vector<std::thread> ts;
    for (uint32_t c = 0; c < clients; ++c) ts.emplace_back(client, srv, infos, c, requests, false);
    for (auto& t : ts) t.join();
    if (blest_server_drain(srv) != BLEST_OK) failf("drain failed %llu %llu", 0, 0);


    // free with requests still queued: the destructor finishes them first
    std::vector<std::unique_ptr<int32_t[]>> late;
    for (uint32_t i = 0; i < 40; ++i) {
        late.emplace_back(new int32_t[infos[1].n]);
        uint64_t t = 0;
        const blest_status st = blest_server_submit(srv, 1, i % 4 ? BLEST_REQ_BFS : BLEST_REQ_CC, i % infos[1].n,
                                                    late.back().get(), &t);
        if (st != BLEST_OK && st != BLEST_ERR_BUSY) failf("late submit %llu status %llu", i, st);
    }
    blest_server_free(srv);

    // The serving benchmark's shape: 4 workers, batches of up to 64 roots (8 or more run
    // as one multi-source launch), CC requests, results written into "device" buffers
    // sized for each graph's n, and a budget of one batch query and two single ones, so
    // idle queries are evicted while other workers run.
    {
        blest_server_opts o2 = o;
        o2.workers = 4;
        o2.max_batch = 64;
        o2.min_multi = 8;
        o2.capacity = 256;
        o2.hbm_budget = mock_footprint(64) + 2 * mock_footprint(1);
        mock_set_budget(o2.hbm_budget);
        blest_server* srv2 = nullptr;
        if (blest_server_create(graphs, 3, &o2, &srv2) != BLEST_OK) {
            std::fprintf(stderr, "server create failed: %s\n", blest_last_error());
            return 1;
        }
        std::vector<std::thread> t2;
        for (uint32_t c = 0; c < clients; ++c) t2.emplace_back(client, srv2, infos, 10 + c, requests, true);
        for (auto& t : t2) t.join();
        if (blest_server_drain(srv2) != BLEST_OK) failf("drain failed %llu %llu", 1, 0);
        blest_server_free(srv2);
    }

    const mock_report r = mock_report_get();
    for (auto* g : graphs) mock_graph_free(g);
    std::printf("queries created %llu, traversal calls %llu, peak bytes %llu (budgets %llu, then %llu), "
                "live after free %llu\n",
                (unsigned long long)r.created, (unsigned long long)r.calls, (unsigned long long)r.peak_bytes,
                (unsigned long long)o.hbm_budget,
                (unsigned long long)(mock_footprint(64) + 2 * mock_footprint(1)), (unsigned long long)r.live);
    if (r.violations || r.live || gFailures.load()) {
        std::printf("server stress FAILED: %llu backend violations, %d result failures\n",
                    (unsigned long long)r.violations, gFailures.load());
        return 1;
    }
    std::printf("server stress ok\n");
    return 0;
}
