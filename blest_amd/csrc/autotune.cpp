// Launch-shape and direction-threshold autotuner for single-source BFS, built on the
// C ABI.  Coordinate descent: segments per CU (a device-graph layout choice, so each
// candidate is a fresh upload), then the top-down slice tile (quads per lane per step)
// and the bottom-up staging depth (words probed together), then alpha, then beta (set
// in place).  Each candidate is timed as the mean HIP-event time over the given roots
// after one warm-up pass, and replaces the best so far only when it is more than 1%
// faster (candidates within timing noise keep the earlier, default-side choice);
// results do not depend on the choice (levels are unique), only the time does.
#include <cstdint>
#include <limits>
#include <memory>
#include <utility>
#include <vector>

#include "blest.h"
#include "capi_internal.hpp"

namespace blest {
namespace {

struct Handles {
    blest_device_graph* dg = nullptr;
    blest_query* q = nullptr;
    ~Handles() {
        if (q) blest_query_free(q);
        if (dg) blest_device_graph_free(dg);
    }
};

void check(blest_status s) {
    if (s != BLEST_OK) throw Error(s, blest_last_error());
}

double time_roots(blest_query* q, const uint32_t* roots, uint32_t nroots) {
    blest_bfs_stats st{};
    for (uint32_t i = 0; i < nroots; ++i) check(blest_bfs(q, roots[i], nullptr, &st));  // warm-up
    double sum = 0;
    for (uint32_t i = 0; i < nroots; ++i) {
        check(blest_bfs(q, roots[i], nullptr, &st));
        sum += st.kernel_ms;
    }
    return sum / nroots;
}

}  // namespace
}  // namespace blest

using namespace blest;

extern "C" blest_status blest_autotune_bfs(const blest_bvss* b, const blest_csc* g, const uint32_t* perm,
                                           const uint32_t* roots, uint32_t nroots, const blest_device_opts* base,
                                           blest_device_opts* best, double* best_ms, uint32_t* trials) {
    return guarded([&] {
        if (!b || !roots || !best) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        if (nroots == 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "the autotuner needs at least one root");
        blest_device_opts o;
        blest_device_opts_default(&o);
        if (base) o = *base;
        if (!g) o.keep_csr = 0;
        if (o.fuse_quads == 0) o.fuse_quads = 2048;
        uint32_t runs = 0;
        constexpr double kMargin = 0.99;  // a new best must be > 1% faster
        double bestMs = std::numeric_limits<double>::infinity();
        blest_device_opts bestOpts = o;
        Handles keep;  // graph of the best layout so far
        for (uint32_t bpc : {1u, 2u, 4u, 8u}) {
            blest_device_opts c = o;
            c.blocks_per_cu = bpc;
            auto h = std::make_unique<Handles>();
            check(blest_device_graph_create(b, g, perm, &c, &h->dg));
            check(blest_query_create(h->dg, 1, &h->q));
            const double ms = time_roots(h->q, roots, nroots);
            ++runs;
            if (ms < bestMs * kMargin) {
                bestMs = ms;
                bestOpts = c;
                std::swap(keep.dg, h->dg);
                std::swap(keep.q, h->q);
            }
        }
        if (keep.dg) {
            for (uint32_t tq : {2u, 1u, 4u})
                for (uint32_t bg : {8u, 4u}) {
                    if (tq == bestOpts.td_quads && bg == bestOpts.bu_group) continue;
                    check(blest_device_graph_set_tiles(keep.dg, tq, bg));
                    const double ms = time_roots(keep.q, roots, nroots);
                    ++runs;
                    if (ms < bestMs * kMargin) {
                        bestMs = ms;
                        bestOpts.td_quads = tq;
                        bestOpts.bu_group = bg;
                    }
                }
            check(blest_device_graph_set_tiles(keep.dg, bestOpts.td_quads, bestOpts.bu_group));
        }
        if (bestOpts.direction == 0 && keep.dg) {
            static const double alphas[] = {4, 8, 15, 30, 60, 120, 250, 500, 1000};
            static const double betas[] = {6, 12, 18, 24, 48, 96, 200, 400, 800};
            for (double a : alphas) {
                if (a == bestOpts.alpha) continue;
                check(blest_device_graph_set_thresholds(keep.dg, a, bestOpts.beta, 0));
                const double ms = time_roots(keep.q, roots, nroots);
                ++runs;
                if (ms < bestMs * kMargin) {
                    bestMs = ms;
                    bestOpts.alpha = a;
                }
            }
            for (double be : betas) {
                if (be == bestOpts.beta) continue;
                check(blest_device_graph_set_thresholds(keep.dg, bestOpts.alpha, be, 0));
                const double ms = time_roots(keep.q, roots, nroots);
                ++runs;
                if (ms < bestMs * kMargin) {
                    bestMs = ms;
                    bestOpts.beta = be;
                }
            }
        }
        // the fused small-level bound (a caller's "never" stays never)
        if (keep.dg && bestOpts.fuse_quads > 0) {
            for (int32_t fq : {1024, 2048, 4096, 512}) {
                if (fq == bestOpts.fuse_quads) continue;
                check(blest_device_graph_set_fuse(keep.dg, fq));
                const double ms = time_roots(keep.q, roots, nroots);
                ++runs;
                if (ms < bestMs * kMargin) {
                    bestMs = ms;
                    bestOpts.fuse_quads = fq;
                }
            }
        }
        *best = bestOpts;
        if (best_ms) *best_ms = bestMs;
        if (trials) *trials = runs;
    });
}
