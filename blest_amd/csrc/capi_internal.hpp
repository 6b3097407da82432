// Helpers shared by the extern "C" translation units.
#pragma once

#include <new>
#include <string>

#include "bvss.hpp"
#include "csc.hpp"

struct blest_csc;
struct blest_bvss;

namespace blest {

extern thread_local std::string tl_last_error;
blest_status fail(blest_status s, const std::string& msg);
const CSC& csc_of(const blest_csc* h);
const BVSS& bvss_of(const blest_bvss* h);
// A new handle owning b.
blest_bvss* bvss_wrap(BVSS&& b);

// Multi-source BFS writing the levels of root j into rows[j] (n entries each); used by
// the request server to hand each request its own row without a staging copy.
// perRoot (nullable, k entries) receives each root's own level count, visited count
// and component edges.
blest_status mbfs_rows(blest_query* q, const uint32_t* srcs, uint32_t k, int32_t* const* rows,
                       blest_bfs_stats* st, blest_bfs_stats* perRoot);

// Marks a query as never replaying captured level chunks (used by the request server,
// whose worker threads allocate while others run).
void query_disable_graphs(blest_query* q);

// Runs fn, converting exceptions into a status code (and the thread's last error).
template <class F>
blest_status guarded(F&& fn) {
    try {
        tl_last_error.clear();
        fn();
        return BLEST_OK;
    } catch (const Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(BLEST_ERR_OUT_OF_HOST_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(BLEST_ERR_INVALID_ARGUMENT, e.what());
    }
}

}  // namespace blest
