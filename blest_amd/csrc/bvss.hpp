// BVSS bit-matrix (host).  Reference: class BVSS : BitMatrix
// (DataStructures/BVSS.cuh:28-595, DataStructures/BitMatrix.cuh:21-30).
//
// Real slice set (rset) r covers columns [r*s, r*s+s) for slice size s.  A slice is a
// row i with at least one nonzero among those columns and an s-bit mask (bit b =
// A(i, r*s+b)); rows ascend within an rset.  The rset's slices are cut into virtual
// slice sets (vsets) of up to 32*noMasks slices laid out for 32 lanes x noMasks
// positions: lane t, position p of a complete vset c holds slice c*128 + p*32 + t,
// stored at rowIds[base + noMasks*t + p] and in mask word masks[base/noMasks + t]
// at bits [s*p, s*p+s) (BVSS.cuh:326-418).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "csc.hpp"

namespace blest {

struct BVSS {
    uint32_t n = 0, sliceSize = 8, noMasks = 4, noRealSliceSets = 0, noVirtualSliceSets = 0;
    bool fullPadding = false;
    uint64_t noSlices = 0, noUnpaddedSlices = 0, noPaddedSlices = 0, edgesCheck = 0;
    double updateDivergence = 0;
    std::vector<uint64_t> sliceSetPtrs;  // V + 1, in slices
    std::vector<uint32_t> virtualToReal;  // V
    std::vector<uint32_t> realPtrs;       // R + 1, in vsets
    std::vector<uint32_t> rowIds;         // S
    std::vector<uint32_t> masks;          // S / noMasks
    // printBVSSData statistics (BVSS.cuh:469-552)
    double avgSlices = 0, stdSlices = 0, compressionRatio = 0;
    uint32_t minSlices = 0, maxSlices = 0;
};

// Throws BLEST_ERR_SLICE_SIZE for the reference's invalid combinations (BVSS.cuh:103-110).
void bvss_check_shape(uint32_t sliceSize, uint32_t noMasks);
BVSS bvss_build(const CSC& g, uint32_t sliceSize, uint32_t noMasks, bool fullPadding,
                unsigned threads = 0);
void bvss_compute_stats(BVSS& b);
// Checks that b's arrays form a BVSS of the reference layout that the device can run:
// prefix-summed sliceSetPtrs / realPtrs, virtualToReal consistent with realPtrs,
// R = ceil(n / sliceSize), vsets of whole mask words, row IDs below n (or the pad row
// 0xFFFFFFFF with an empty mask).  Throws BLEST_ERR_INVALID_ARGUMENT naming the fault.
void bvss_validate(const BVSS& b);
void bvss_save_binary(const BVSS& b, const std::string& path);  // BVSS.cuh:125-146
BVSS bvss_load_binary(const std::string& path);                 // BVSS.cuh:148-183
std::string bvss_csv_header();
std::string bvss_csv_row(const BVSS& b);

}  // namespace blest
