// CSC adjacency pattern and the host-side graph utilities around it.
// Reference type: class CSC (DataStructures/CSC.cuh:35-96): column j lists the rows i
// of the edges j -> i, every column sorted ascending.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace blest {

struct CSC {
    uint32_t n = 0;
    std::vector<uint32_t> colPtrs;  // n + 1 (u32, as CSC::m_ColPtrs)
    std::vector<uint32_t> rows;     // nnz
    bool isSocial = false;          // CSC::m_IsSocial (CSC.cuh:176)

    uint64_t nnz() const { return rows.size(); }
    uint32_t degree(uint32_t j) const { return colPtrs[j + 1] - colPtrs[j]; }
};

// Sort every column, drop self-loops and duplicate rows, recompute isSocial.
void csc_normalise(CSC& g, unsigned threads = 0);
// Build from a caller-given pattern (copied; then normalised).
CSC csc_from_arrays(uint32_t n, const uint32_t* colPtrs, const uint32_t* rows);
// CSC::CSC(filename, undirected, binary) (CSC.cuh:98-185).
CSC csc_from_mtx(const std::string& path, bool undirected, bool binary);
// Binary cache (CSC.cuh:225-266).
CSC csc_load_binary(const std::string& path);
void csc_save_binary(const CSC& g, const std::string& path);

// Symmetric graph from an edge stream: edge(k, &u, &v) for k in [0, m) returns false
// to skip a sample.  Self-loops are dropped, both directions stored, duplicates merged.
using EdgeFn = bool (*)(const void* ctx, uint64_t k, uint32_t* u, uint32_t* v);
CSC csc_from_edge_stream(uint32_t n, uint64_t m, EdgeFn fn, const void* ctx, unsigned threads);

CSC generate_rmat(uint32_t scale, uint32_t edgeFactor, double a, double b, double c,
                  uint64_t seed, bool permute, unsigned threads = 0);
CSC generate_grid(uint32_t side, double keep, double shortcutFrac, uint64_t seed,
                  unsigned threads = 0);
CSC generate_er(uint32_t n, double avgDegree, uint64_t seed, unsigned threads = 0);
// Seeded Fisher-Yates permutation of [0, n): perm[i] = image of i.
std::vector<uint32_t> seeded_permutation(uint32_t n, uint64_t seed, uint64_t stream);

// CSC::transpose (CSC.cuh:437-476).
CSC csc_transpose(const CSC& g, unsigned threads = 0);
// CSC::symmetrize (CSC.cuh:478-523).
CSC csc_symmetrize(const CSC& g, unsigned threads = 0);
// CSC::checkSymmetry (CSC.cuh:410-435); binary-search variant.
bool csc_is_symmetric(const CSC& g, unsigned threads = 0);
// isSocialDegreeDistribution (Common.cuh:100-250) on in- and out-degree, OR-ed with
// averageDegree > SOCIAL_THRESHOLD (CSC.cuh:176,187-217).
bool csc_is_social(const CSC& g);
bool social_degree_distribution(const std::vector<uint32_t>& deg);
// CSC::applyPermutation (CSC.cuh:1144-1191); throws BLEST_ERR_PERMUTATION when
// invPerm is not a bijection (CSC.cuh:1193-1210).
void csc_apply_permutation(CSC& g, const uint32_t* invPerm, unsigned threads = 0);
bool permutation_is_bijection(const uint32_t* invPerm, uint32_t n);

struct CSCMetrics {
    uint32_t maxBandwidth, maxProfile, maxDegree;
    double avgBandwidth, avgProfile, avgDegree;
};
CSCMetrics csc_metrics(const CSC& g);  // CSC.cuh:268-333

std::vector<uint32_t> pick_roots(const CSC& g, uint32_t count, uint64_t seed);

}  // namespace blest
