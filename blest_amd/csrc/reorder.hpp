// Host-side vertex reorderings that precede BVSS construction.
// Reference dispatch: CSC::reorder (DataStructures/CSC.cuh:358-386).
#pragma once

#include <vector>

#include "csc.hpp"

namespace blest {

// Each function returns the inverse permutation invPerm[old] = new; none of them
// modifies g.  reorder() applies the chosen permutation to g.
std::vector<uint32_t> order_natural(const CSC& g);
std::vector<uint32_t> order_random(const CSC& g, uint64_t seed);
std::vector<uint32_t> order_degree(const CSC& g, bool ascending);
std::vector<uint32_t> order_rcm(const CSC& g, uint64_t seed);
std::vector<uint32_t> order_jaccard_window(const CSC& g, uint32_t sliceSize, uint32_t window,
                                           unsigned threads);
// threads > 1: the greedy's priority updates spread over up to 8 threads (the same order
// as with one thread).
std::vector<uint32_t> order_gorder(const CSC& g, uint32_t window, unsigned threads = 1);

// chainPermutations (Common.cuh:81-91): result[i] = second[first[i]].
std::vector<uint32_t> chain_permutations(const std::vector<uint32_t>& first,
                                         const std::vector<uint32_t>& second);

// Reorders g in place and returns invPerm.
std::vector<uint32_t> reorder(CSC& g, const blest_reorder_opts& o);

}  // namespace blest
