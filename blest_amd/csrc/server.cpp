// Request server over resident device graphs: continuous batching of single-root BFS
// and CC requests onto a few concurrent HIP streams, bounded by an HBM byte budget.
//
// Built only on the C ABI (device graphs, queries, blest_bfs / blest_cc) plus the
// row-scatter multi-source run.  Each worker thread owns one query at a time, and a
// query owns its HIP stream, so batches of different workers overlap on the device.
//
// Scheduling: the queue is FIFO.  A free worker takes the oldest request; when it is a
// BFS it also takes every other queued BFS on the same graph (oldest first, up to
// max_batch) and runs them as one multi-source traversal.  BFS levels and canonical CC
// labels are unique, so a request's result does not depend on its batch.
//
// Memory: queries are pooled per (graph, capacity) and counted against the budget with
// blest_query_footprint.  A worker that needs a query for which the budget has no room
// frees idle pooled queries, least recently used first, and otherwise waits for a
// running batch to return its query.  Any single query must fit the budget.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "blest.h"
#include "capi_internal.hpp"

using Clock = std::chrono::steady_clock;

namespace blest {
namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t ticket = 0;
    uint32_t graph = 0;
    blest_request_kind kind = BLEST_REQ_BFS;
    uint32_t root = 0;
    void* out = nullptr;
    Clock::time_point submitted;
    bool done = false;
    blest_request_result res{};
    std::string error;
};

struct PooledQuery {
    blest_query* q = nullptr;
    uint32_t graph = 0;
    uint32_t capacity = 0;  // max_sources
    uint64_t bytes = 0;
    bool busy = false;
    uint64_t lastUse = 0;
};

}  // namespace

struct Server {
    std::vector<blest_device_graph*> graphs;
    std::vector<blest_device_graph_info> info;
    blest_server_opts opts{};

    std::mutex mu;
    std::condition_variable cvWork;   // workers: a request was queued / stop
    std::condition_variable cvDone;   // waiters: a request completed
    std::condition_variable cvPool;   // workers: a pooled query was returned
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
    uint64_t nextTicket = 1;
    uint64_t outstanding = 0;  // queued + running
    bool stop = false;

    std::list<PooledQuery> pool;  // guarded by mu
    uint64_t poolBytes = 0;
    uint64_t useClock = 0;

    std::vector<std::thread> workers;

    ~Server() {
        {
            std::unique_lock<std::mutex> lk(mu);
            cvDone.wait(lk, [&] { return outstanding == 0; });
            stop = true;
        }
        cvWork.notify_all();
        for (auto& t : workers) t.join();
        for (auto& p : pool) blest_query_free(p.q);
    }

    // Takes the next batch (oldest request first); empty when stopping.
    std::vector<std::shared_ptr<Request>> take_batch(std::unique_lock<std::mutex>& lk) {
        cvWork.wait(lk, [&] { return stop || !queue.empty(); });
        std::vector<std::shared_ptr<Request>> batch;
        if (queue.empty()) return batch;
        batch.push_back(queue.front());
        queue.pop_front();
        const Request& head = *batch[0];
        if (head.kind == BLEST_REQ_BFS && opts.max_batch > 1) {
            for (auto it = queue.begin(); it != queue.end() && batch.size() < opts.max_batch;) {
                if ((*it)->kind == BLEST_REQ_BFS && (*it)->graph == head.graph) {
                    batch.push_back(*it);
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
        }
        return batch;
    }

    // Checks out a pooled query for (graph, >= capacity), creating one within the budget.
    PooledQuery* acquire(std::unique_lock<std::mutex>& lk, uint32_t graph, uint32_t capacity) {
        for (;;) {
            PooledQuery* best = nullptr;
            for (auto& p : pool)
                if (!p.busy && p.graph == graph && p.capacity >= capacity && (!best || p.capacity < best->capacity))
                    best = &p;
            if (best) {
                best->busy = true;
                best->lastUse = ++useClock;
                return best;
            }
            uint64_t need = 0;
            blest_status st = blest_query_footprint(graphs[graph], capacity, &need);
            if (st != BLEST_OK) throw Error(st, blest_last_error());
            if (need > opts.hbm_budget)
                throw Error(BLEST_ERR_OUT_OF_DEVICE_MEMORY, "one query of this graph exceeds the server's HBM budget");
            // free idle queries, least recently used first, until the new one fits
            while (poolBytes + need > opts.hbm_budget) {
                auto victim = pool.end();
                for (auto it = pool.begin(); it != pool.end(); ++it)
                    if (!it->busy && (victim == pool.end() || it->lastUse < victim->lastUse)) victim = it;
                if (victim == pool.end()) break;
                blest_query_free(victim->q);
                poolBytes -= victim->bytes;
                pool.erase(victim);
            }
            if (poolBytes + need > opts.hbm_budget) {
                cvPool.wait(lk);  // every remaining query is running; one will come back
                continue;
            }
            poolBytes += need;  // reserve before dropping the lock for the allocation
            lk.unlock();
            blest_query* q = nullptr;
            st = blest_query_create(graphs[graph], capacity, &q);
            std::string why = st == BLEST_OK ? std::string() : std::string(blest_last_error());
            lk.lock();
            if (st != BLEST_OK) {
                poolBytes -= need;
                cvPool.notify_all();
                throw Error(st, why);
            }
            query_disable_graphs(q);
            pool.push_back(PooledQuery{q, graph, capacity, need, true, ++useClock});
            return &pool.back();
        }
    }

    void release(PooledQuery* p) {
        p->busy = false;
        cvPool.notify_all();
    }

    void run_batch(std::vector<std::shared_ptr<Request>>& batch) {
        const uint32_t g = batch[0]->graph;
        const uint32_t k = uint32_t(batch.size());
        const bool cc = batch[0]->kind == BLEST_REQ_CC;
        const Clock::time_point start = Clock::now();
        blest_status st = BLEST_OK;
        std::string why;
        blest_bfs_stats bs{};
        std::vector<blest_bfs_stats> perRoot(k);
        blest_cc_stats cs{};
        PooledQuery* pq = nullptr;
        const bool multi = !cc && k > 1 && k >= opts.min_multi;
        try {
            std::unique_lock<std::mutex> lk(mu);
            // single-source runs use single-source state; a multi-source launch the batch capacity
            pq = acquire(lk, g, multi ? opts.max_batch : 1);
        } catch (const Error& e) {
            st = e.code;
            why = e.what();
        }
        if (pq) {
            if (cc) {
                st = blest_cc(pq->q, 0, static_cast<uint32_t*>(batch[0]->out), &cs);
            } else if (!multi) {
                for (uint32_t j = 0; j < k && st == BLEST_OK; ++j) {
                    st = blest_bfs(pq->q, batch[j]->root, static_cast<int32_t*>(batch[j]->out), &perRoot[j]);
                    bs = perRoot[j];
                }
            } else {
                std::vector<uint32_t> roots(k);
                std::vector<int32_t*> rows(k);
                for (uint32_t j = 0; j < k; ++j) {
                    roots[j] = batch[j]->root;
                    rows[j] = static_cast<int32_t*>(batch[j]->out);
                }
                st = mbfs_rows(pq->q, roots.data(), k, rows.data(), &bs, perRoot.data());
            }
            if (st != BLEST_OK) why = blest_last_error();
        }
        const Clock::time_point end = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        if (pq) release(pq);
        for (uint32_t j = 0; j < k; ++j) {
            Request* r = batch[j].get();
            blest_request_result& res = r->res;
            res.ticket = r->ticket;
            res.status = st;
            res.batch_size = k;
            res.queued_ms = ms_between(r->submitted, start);
            res.service_ms = ms_between(start, end);
            if (cc) {
                res.kernel_ms = cs.kernel_ms;
                res.total_levels = cs.iterations;
                res.no_visited = cs.no_components;
                res.edges_traversed_component = 0;
            } else {
                res.kernel_ms = perRoot[j].kernel_ms;
                res.total_levels = perRoot[j].total_levels;
                res.no_visited = perRoot[j].no_visited;
                res.edges_traversed_component = perRoot[j].edges_traversed_component;
            }
            r->error = why;
            r->done = true;
            --outstanding;
        }
        cvDone.notify_all();
    }

    void worker_loop() {
        for (;;) {
            std::vector<std::shared_ptr<Request>> batch;
            {
                std::unique_lock<std::mutex> lk(mu);
                batch = take_batch(lk);
            }
            if (batch.empty()) return;
            run_batch(batch);
        }
    }
};

}  // namespace blest

struct blest_server {
    blest::Server s;
};

using namespace blest;

extern "C" void blest_server_opts_default(blest_server_opts* o) {
    if (!o) return;
    o->workers = 4;
    o->max_batch = 64;
    o->hbm_budget = 0;
    o->capacity = 65536;
    o->min_multi = 8;
}

extern "C" blest_status blest_server_create(blest_device_graph* const* graphs, uint32_t count,
                                            const blest_server_opts* o, blest_server** out) {
    return guarded([&] {
        if (!out || (count && !graphs)) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if (count == 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "a server needs at least one graph");
        blest_server_opts opts;
        blest_server_opts_default(&opts);
        if (o) opts = *o;
        if (opts.workers == 0 || opts.workers > 64) throw Error(BLEST_ERR_INVALID_ARGUMENT, "workers must be 1..64");
        if (opts.max_batch == 0 || opts.max_batch > 128)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "max_batch must be 1..128");
        if (opts.capacity == 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "capacity must be positive");
        auto h = std::make_unique<blest_server>();
        Server& s = h->s;
        s.graphs.assign(graphs, graphs + count);
        s.info.resize(count);
        for (uint32_t i = 0; i < count; ++i) {
            if (!graphs[i]) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null device graph");
            blest_status st = blest_device_graph_info_get(graphs[i], &s.info[i]);
            if (st != BLEST_OK) throw Error(st, blest_last_error());
        }
        for (uint32_t i = 1; i < count; ++i)
            if (s.info[i].device != s.info[0].device)
                throw Error(BLEST_ERR_INVALID_ARGUMENT, "all graphs of a server live on one device");
        if (opts.hbm_budget == 0) {
            uint64_t freeB = 0, totalB = 0;
            blest_status st = blest_device_memory_info(s.info[0].device, &freeB, &totalB);
            if (st != BLEST_OK) throw Error(st, blest_last_error());
            opts.hbm_budget = freeB / 10 * 8;
        }
        s.opts = opts;
        for (uint32_t w = 0; w < opts.workers; ++w) s.workers.emplace_back([&s] { s.worker_loop(); });
        *out = h.release();
    });
}

extern "C" blest_status blest_server_submit(blest_server* h, uint32_t graph, blest_request_kind kind, uint32_t root,
                                            void* out, uint64_t* ticket) {
    return guarded([&] {
        if (!h || !out || !ticket) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Server& s = h->s;
        if (graph >= s.graphs.size()) throw Error(BLEST_ERR_INVALID_ARGUMENT, "graph index out of range");
        if (kind != BLEST_REQ_BFS && kind != BLEST_REQ_CC) throw Error(BLEST_ERR_INVALID_ARGUMENT, "unknown request kind");
        if (kind == BLEST_REQ_BFS && root >= s.info[graph].n)
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "source vertex out of range");
        if (kind == BLEST_REQ_CC && !s.info[graph].symmetric)
            throw Error(BLEST_ERR_UNSUPPORTED, "connected components need a symmetric graph");
        auto r = std::make_shared<Request>();
        r->graph = graph;
        r->kind = kind;
        r->root = root;
        r->out = out;
        {
            std::lock_guard<std::mutex> lk(s.mu);
            if (s.stop) throw Error(BLEST_ERR_INVALID_ARGUMENT, "server is shutting down");
            if (s.outstanding >= s.opts.capacity) throw Error(BLEST_ERR_BUSY, "request queue is full");
            r->ticket = s.nextTicket++;
            r->submitted = Clock::now();
            s.tickets.emplace(r->ticket, r);
            s.queue.push_back(r);
            ++s.outstanding;
            *ticket = r->ticket;
        }
        s.cvWork.notify_one();
    });
}

extern "C" blest_status blest_server_wait(blest_server* h, uint64_t ticket, uint32_t timeout_ms,
                                          blest_request_result* res) {
    return guarded([&] {
        if (!h) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Server& s = h->s;
        std::unique_lock<std::mutex> lk(s.mu);
        auto it = s.tickets.find(ticket);
        if (it == s.tickets.end()) throw Error(BLEST_ERR_INVALID_ARGUMENT, "unknown or already collected ticket");
        std::shared_ptr<Request> r = it->second;
        auto ready = [&] { return r->done; };
        if (timeout_ms == 0xFFFFFFFFu)
            s.cvDone.wait(lk, ready);
        else if (!s.cvDone.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
            throw Error(BLEST_ERR_TIMEOUT, "request still running");
        s.tickets.erase(r->ticket);
        if (res) *res = r->res;
        if (r->res.status != BLEST_OK)
            throw Error(static_cast<blest_status>(r->res.status), "request failed: " + r->error);
    });
}

extern "C" blest_status blest_server_drain(blest_server* h) {
    return guarded([&] {
        if (!h) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Server& s = h->s;
        std::unique_lock<std::mutex> lk(s.mu);
        s.cvDone.wait(lk, [&] { return s.outstanding == 0; });
    });
}

extern "C" void blest_server_free(blest_server* h) { delete h; }
