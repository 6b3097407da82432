// Multi-source BFS and the many-graphs batch (placeholders until their kernels land).
#include "dgraph.hpp"

using namespace blest;

extern "C" {

blest_status blest_mbfs(blest_query* qh, const uint32_t* srcs, uint32_t k, int32_t* levels,
                        const blest_mbfs_opts* o, blest_bfs_stats* st) {
    (void)qh;
    (void)srcs;
    (void)k;
    (void)levels;
    (void)o;
    (void)st;
    return fail(BLEST_ERR_UNSUPPORTED, "multi-source BFS not built yet");
}

blest_status blest_graph_batch_create(const blest_csc* const* graphs, uint32_t count, int device,
                                      blest_graph_batch** out) {
    (void)graphs;
    (void)count;
    (void)device;
    (void)out;
    return fail(BLEST_ERR_UNSUPPORTED, "graph batches not built yet");
}

void blest_graph_batch_free(blest_graph_batch* gb) { (void)gb; }

blest_status blest_graph_batch_bfs(blest_graph_batch* gb, const uint32_t* roots, int32_t* levels,
                                   double* kernel_ms) {
    (void)gb;
    (void)roots;
    (void)levels;
    (void)kernel_ms;
    return fail(BLEST_ERR_UNSUPPORTED, "graph batches not built yet");
}

}  // extern "C"
