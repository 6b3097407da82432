// C entry point over the reference's vendored Gorder (Gorder/Graph.cpp, MIT, (c) 2016
// Hao Wei), compiled from the reference tree into oracle/_ref/ by oracle/Makefile.
// TEST INFRASTRUCTURE ONLY: used by tests/ to compare this build's Gorder ordering
// against the reference's on small graphs.  Mirrors CSC::gorderWithJackard's calls
// (DataStructures/CSC.cuh:934-942): readGraph, Transform, GorderGreedy.
#include <cstdint>
#include <vector>

#include "Graph.h"

extern "C" int gorder_ref(uint32_t n, uint32_t nnz, uint32_t* col_ptrs, uint32_t* rows, long long window,
                          int transform, int64_t* order_out) {
    Gorder::Graph g;
    g.setFilename("gorder");
    g.readGraph(n, nnz, col_ptrs, rows);
    if (transform) g.Transform();
    std::vector<long long int> order;
    g.GorderGreedy(order, window);
    for (uint32_t i = 0; i < n; ++i) order_out[i] = order[i];
    return 0;
}
