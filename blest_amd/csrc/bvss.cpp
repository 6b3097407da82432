// BVSS construction from CSC (host).
#include "bvss.hpp"

#include <bit>
#include <cmath>
#include <fstream>
#include <sstream>

namespace blest {

namespace {
constexpr uint32_t kLanes = 32;  // lanes per vset in the reference layout (WARP_SIZE)
constexpr uint32_t kK = 128;     // Common.cuh:40

// Number of vsets and total slot count (slices incl. padding) an rset of `cnt`
// slices occupies (BVSS.cuh:326-418).
inline void rset_shape(uint64_t cnt, uint32_t noMasks, bool fullPadding, uint32_t& vsets,
                       uint64_t& slots) {
    const uint64_t full = uint64_t(kLanes) * noMasks;
    const uint64_t complete = cnt / full;
    const uint64_t left = cnt - complete * full;
    vsets = uint32_t(complete + (left ? 1 : 0));
    slots = complete * full;
    if (left) {
        if (fullPadding) {
            slots += full;
        } else {
            const uint64_t stride = left / noMasks;
            slots += stride * noMasks + ((left - stride * noMasks) ? noMasks : 0);
        }
    }
}

// Size in slots of vset `v` (0-based within its rset).
inline uint64_t vset_slots(uint64_t cnt, uint32_t noMasks, bool fullPadding, uint32_t v) {
    const uint64_t full = uint64_t(kLanes) * noMasks;
    const uint64_t complete = cnt / full;
    if (v < complete) return full;
    const uint64_t left = cnt - complete * full;
    if (fullPadding) return full;
    const uint64_t stride = left / noMasks;
    return stride * noMasks + ((left - stride * noMasks) ? noMasks : 0);
}

// Merge the (sorted) columns of rset r into ascending (row, mask) pairs.  Calls
// emit(row, mask) per slice; returns the slice count.
template <class Emit>
uint64_t merge_rset(const CSC& g, uint32_t r, uint32_t s, Emit&& emit) {
    const uint32_t c0 = r * s, c1 = std::min<uint64_t>(uint64_t(c0) + s, g.n);
    const uint32_t k = c1 - c0;
    uint32_t pos[32], end[32];
    for (uint32_t b = 0; b < k; ++b) {
        pos[b] = g.colPtrs[c0 + b];
        end[b] = g.colPtrs[c0 + b + 1];
    }
    uint64_t count = 0;
    for (;;) {
        uint32_t row = kNoVertex;
        for (uint32_t b = 0; b < k; ++b)
            if (pos[b] < end[b] && g.rows[pos[b]] < row) row = g.rows[pos[b]];
        if (row == kNoVertex) break;
        uint32_t mask = 0;
        for (uint32_t b = 0; b < k; ++b)
            while (pos[b] < end[b] && g.rows[pos[b]] == row) {  // duplicates OR together
                mask |= 1u << b;
                ++pos[b];
            }
        emit(row, mask);
        ++count;
    }
    return count;
}
}  // namespace

void bvss_check_shape(uint32_t sliceSize, uint32_t noMasks) {
    if (sliceSize == 0 || sliceSize > 32 || kK % sliceSize != 0)
        throw Error(BLEST_ERR_SLICE_SIZE, "Invalid slice size provided.");
    if (noMasks != 32 / sliceSize) throw Error(BLEST_ERR_SLICE_SIZE, "Invalid noMasks provided.");
}

BVSS bvss_build(const CSC& g, uint32_t sliceSize, uint32_t noMasks, bool fullPadding,
                unsigned threads) {
    bvss_check_shape(sliceSize, noMasks);
    threads = hw_threads(threads);
    BVSS b;
    b.n = g.n;
    b.sliceSize = sliceSize;
    b.noMasks = noMasks;
    b.fullPadding = fullPadding;
    const uint32_t R = uint32_t((uint64_t(g.n) + sliceSize - 1) / sliceSize);
    b.noRealSliceSets = R;

    // Pass 1: slices per rset -> vset counts and slot counts.
    std::vector<uint64_t> cnt(R);
    parallel_dynamic(R, threads, 1024, [&](uint64_t r, unsigned) {
        cnt[r] = merge_rset(g, uint32_t(r), sliceSize, [](uint32_t, uint32_t) {});
    });
    b.realPtrs.assign(size_t(R) + 1, 0);
    std::vector<uint64_t> rsetSlot(size_t(R) + 1, 0);
    for (uint32_t r = 0; r < R; ++r) {
        uint32_t v;
        uint64_t slots;
        rset_shape(cnt[r], noMasks, fullPadding, v, slots);
        b.realPtrs[r + 1] = b.realPtrs[r] + v;
        rsetSlot[r + 1] = rsetSlot[r] + slots;
    }
    const uint32_t V = b.realPtrs[R];
    b.noVirtualSliceSets = V;
    b.noSlices = rsetSlot[R];
    b.sliceSetPtrs.assign(size_t(V) + 1, 0);
    b.virtualToReal.resize(V);
    b.rowIds.resize(b.noSlices);
    b.masks.assign(b.noSlices / noMasks, 0);
    std::vector<double> vsetDiv(V, 0.0);

    // Pass 2: lay out each rset's vsets.
    parallel_dynamic(R, threads, 512, [&](uint64_t r, unsigned) {
        const uint64_t c = cnt[r];
        std::vector<uint32_t> rows;
        std::vector<uint32_t> msk;
        rows.reserve(c);
        msk.reserve(c);
        merge_rset(g, uint32_t(r), sliceSize, [&](uint32_t row, uint32_t m) {
            rows.push_back(row);
            msk.push_back(m);
        });
        uint64_t slot = rsetSlot[r];
        const uint32_t v0 = b.realPtrs[r], v1 = b.realPtrs[r + 1];
        const uint64_t full = uint64_t(kLanes) * noMasks;
        const uint64_t complete = c / full;
        for (uint32_t v = v0; v < v1; ++v) {
            const uint32_t vi = v - v0;
            const uint64_t size = vset_slots(c, noMasks, fullPadding, vi);
            b.sliceSetPtrs[v + 1] = slot + size;  // prefix filled per vset
            b.virtualToReal[v] = uint32_t(r);
            uint32_t* outRows = b.rowIds.data() + slot;
            uint32_t* outMasks = b.masks.data() + slot / noMasks;
            auto put = [&](uint64_t lane, uint32_t p, uint64_t src, bool real) {
                outRows[lane * noMasks + p] = real ? rows[src] : kNoVertex;
                if (real) outMasks[lane] |= msk[src] << (sliceSize * p);
            };
            if (vi < complete) {  // BVSS.cuh:330-348
                for (uint32_t t = 0; t < kLanes; ++t)
                    for (uint32_t p = 0; p < noMasks; ++p) put(t, p, vi * full + uint64_t(p) * kLanes + t, true);
            } else if (fullPadding) {  // BVSS.cuh:356-378
                const uint64_t ls = complete * full;
                for (uint32_t t = 0; t < kLanes; ++t)
                    for (uint32_t p = 0; p < noMasks; ++p) {
                        uint64_t cur = ls + uint64_t(p) * kLanes + t;
                        put(t, p, cur, cur < c);
                    }
            } else {  // BVSS.cuh:379-415
                uint64_t ls = complete * full;
                const uint64_t stride = (c - ls) / noMasks;
                for (uint64_t t = 0; t < stride; ++t)
                    for (uint32_t p = 0; p < noMasks; ++p) put(t, p, ls + uint64_t(p) * stride + t, true);
                ls += stride * noMasks;
                if (ls < c)
                    for (uint32_t p = 0; p < noMasks; ++p) put(stride, p, ls + p, ls + p < c);
            }
            // computeUpdateDivergence term of this vset (BVSS.cuh:420-467): mean over
            // positions of the population std-dev of the rows the 32 lanes hold there.
            double vmean = 0;
            for (uint32_t p = 0; p < noMasks; ++p) {
                double mean = 0, sd = 0;
                uint32_t k = 0;
                for (uint32_t t = 0; t < kLanes; ++t) {
                    uint64_t sl = uint64_t(t) * noMasks + p;
                    if (sl < size && outRows[sl] != kNoVertex) {
                        mean += outRows[sl];
                        ++k;
                    }
                }
                if (k) mean /= k;
                for (uint32_t t = 0; t < kLanes; ++t) {
                    uint64_t sl = uint64_t(t) * noMasks + p;
                    if (sl < size && outRows[sl] != kNoVertex) sd += (outRows[sl] - mean) * (outRows[sl] - mean);
                }
                if (k) sd /= k;
                vmean += std::sqrt(sd);
            }
            vsetDiv[v] = vmean / noMasks;
            slot += size;
        }
    });
    double div = 0;
    for (uint32_t v = 0; v < V; ++v) div += vsetDiv[v];
    b.updateDivergence = V ? div / V : 0.0;
    bvss_compute_stats(b);
    return b;
}

void bvss_compute_stats(BVSS& b) {
    const uint32_t V = b.noVirtualSliceSets;
    double avg = 0;
    uint32_t mn = kNoVertex, mx = 0;
    for (uint32_t v = 0; v < V; ++v) {
        uint32_t s = uint32_t(b.sliceSetPtrs[v + 1] - b.sliceSetPtrs[v]);
        avg += s;
        mn = std::min(mn, s);
        mx = std::max(mx, s);
    }
    avg = V ? avg / V : 0;
    double var = 0;
    for (uint32_t v = 0; v < V; ++v) {
        double d = double(b.sliceSetPtrs[v + 1] - b.sliceSetPtrs[v]) - avg;
        var += d * d;
    }
    b.avgSlices = avg;
    b.stdSlices = V ? std::sqrt(var / V) : 0;
    b.minSlices = V ? mn : 0;
    b.maxSlices = mx;
    const uint32_t conceal = b.sliceSize == 32 ? 0xFFFFFFFFu : ((1u << b.sliceSize) - 1);
    uint64_t padded = 0, edges = 0;
    for (uint64_t w = 0; w < b.masks.size(); ++w)
        for (uint32_t p = 0; p < b.noMasks; ++p) {
            uint32_t pat = (b.masks[w] >> (p * b.sliceSize)) & conceal;
            uint32_t e = uint32_t(std::popcount(pat));
            padded += e == 0;
            edges += e;
        }
    b.noPaddedSlices = padded;
    b.noUnpaddedSlices = b.noSlices - padded;
    b.edgesCheck = edges;
    b.compressionRatio = b.noUnpaddedSlices ? (double(edges) / double(b.noUnpaddedSlices)) / b.sliceSize : 0;
}

void bvss_validate(const BVSS& b) {
    auto bad = [](const std::string& m) { throw Error(BLEST_ERR_INVALID_ARGUMENT, "BVSS arrays: " + m); };
    bvss_check_shape(b.sliceSize, b.noMasks);
    const uint64_t R = b.noRealSliceSets, V = b.noVirtualSliceSets;
    if (R != (uint64_t(b.n) + b.sliceSize - 1) / b.sliceSize) bad("noRealSliceSets != ceil(n / sliceSize)");
    if (b.sliceSetPtrs.size() != V + 1 || b.virtualToReal.size() != V || b.realPtrs.size() != R + 1)
        bad("array lengths disagree with the slice-set counts");
    if (b.noSlices % b.noMasks || b.rowIds.size() != b.noSlices || b.masks.size() != b.noSlices / b.noMasks)
        bad("noSlices is not a whole number of mask words");
    if (b.sliceSetPtrs[0] != 0 || b.sliceSetPtrs[V] != b.noSlices) bad("sliceSetPtrs must run from 0 to noSlices");
    for (uint64_t v = 0; v < V; ++v) {
        if (b.sliceSetPtrs[v + 1] < b.sliceSetPtrs[v]) bad("sliceSetPtrs decreases at vset " + std::to_string(v));
        if (b.sliceSetPtrs[v] % b.noMasks) bad("vset " + std::to_string(v) + " does not start on a mask word");
    }
    if (b.realPtrs[0] != 0 || b.realPtrs[R] != V) bad("realPtrs must run from 0 to noVirtualSliceSets");
    for (uint64_t r = 0; r < R; ++r) {
        if (b.realPtrs[r + 1] < b.realPtrs[r]) bad("realPtrs decreases at rset " + std::to_string(r));
        for (uint32_t v = b.realPtrs[r]; v < b.realPtrs[r + 1]; ++v)
            if (b.virtualToReal[v] != r) bad("virtualToReal[" + std::to_string(v) + "] != its rset");
    }
    const uint32_t conceal = b.sliceSize == 32 ? 0xFFFFFFFFu : ((1u << b.sliceSize) - 1);
    for (uint64_t sl = 0; sl < b.noSlices; ++sl) {
        const uint32_t m = (b.masks[sl / b.noMasks] >> (b.sliceSize * (sl % b.noMasks))) & conceal;
        const uint32_t row = b.rowIds[sl];
        if (row == kNoVertex ? m != 0 : row >= b.n) bad("row id out of range at slice " + std::to_string(sl));
    }
}

void bvss_save_binary(const BVSS& b, const std::string& path) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw Error(BLEST_ERR_IO, "cannot write BVSS cache " + path);
    auto w = [&](const void* p, size_t sz) { out.write(static_cast<const char*>(p), std::streamsize(sz)); };
    w(&b.n, 4);
    w(&b.sliceSize, 4);
    w(&b.noMasks, 4);
    w(&b.noRealSliceSets, 4);
    w(&b.noVirtualSliceSets, 4);
    w(&b.noSlices, 8);
    w(&b.noUnpaddedSlices, 8);
    w(&b.noPaddedSlices, 8);
    w(&b.updateDivergence, 8);
    w(b.sliceSetPtrs.data(), 8 * b.sliceSetPtrs.size());
    w(b.virtualToReal.data(), 4 * b.virtualToReal.size());
    w(b.realPtrs.data(), 4 * b.realPtrs.size());
    w(b.rowIds.data(), 4 * b.rowIds.size());
    w(b.masks.data(), 4 * b.masks.size());
    if (!out) throw Error(BLEST_ERR_IO, "write failed " + path);
}

BVSS bvss_load_binary(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error(BLEST_ERR_IO, "cannot open BVSS cache " + path);
    BVSS b;
    auto r = [&](void* p, size_t sz) { in.read(static_cast<char*>(p), std::streamsize(sz)); };
    r(&b.n, 4);
    r(&b.sliceSize, 4);
    r(&b.noMasks, 4);
    r(&b.noRealSliceSets, 4);
    r(&b.noVirtualSliceSets, 4);
    r(&b.noSlices, 8);
    r(&b.noUnpaddedSlices, 8);
    r(&b.noPaddedSlices, 8);
    r(&b.updateDivergence, 8);
    if (!in) throw Error(BLEST_ERR_IO, "truncated BVSS cache " + path);
    bvss_check_shape(b.sliceSize, b.noMasks);
    b.sliceSetPtrs.resize(size_t(b.noVirtualSliceSets) + 1);
    b.virtualToReal.resize(b.noVirtualSliceSets);
    b.realPtrs.resize(size_t(b.noRealSliceSets) + 1);
    b.rowIds.resize(b.noSlices);
    b.masks.resize(b.noSlices / b.noMasks);
    r(b.sliceSetPtrs.data(), 8 * b.sliceSetPtrs.size());
    r(b.virtualToReal.data(), 4 * b.virtualToReal.size());
    r(b.realPtrs.data(), 4 * b.realPtrs.size());
    r(b.rowIds.data(), 4 * b.rowIds.size());
    r(b.masks.data(), 4 * b.masks.size());
    if (!in) throw Error(BLEST_ERR_IO, "truncated BVSS cache " + path);
    bvss_validate(b);
    // The cache does not record FULL_PADDING: a padded leftover vset has 32*noMasks slots.
    b.fullPadding = false;
    for (uint32_t v = 0; v < b.noVirtualSliceSets; ++v) {
        uint64_t s = b.sliceSetPtrs[v + 1] - b.sliceSetPtrs[v];
        if (s == uint64_t(kLanes) * b.noMasks) b.fullPadding = true;
    }
    bvss_compute_stats(b);
    return b;
}

std::string bvss_csv_header() {
    return "N\tNNZ\tUpdateDivergence\tSliceSize\t#RSet\t#VSet\t#Slices\tAvg(Slice/VSet)\tMin(Slice/VSet)\t"
           "Max(Slice/VSet)\tStd(Slice/Vset)\t#PaddedSlices\tUnpaddedSlices\tBitsTotal\tBitsUnpadded\t"
           "CompressionRatio";
}

std::string bvss_csv_row(const BVSS& b) {
    std::ostringstream o;
    auto cell = [&](auto v) { o << "'" << v << '\t'; };  // fileFlush (Common.cuh:75-79)
    cell(b.n);
    cell(b.edgesCheck);
    cell(b.updateDivergence);
    cell(b.sliceSize);
    cell(b.noRealSliceSets);
    cell(b.noVirtualSliceSets);
    cell(b.noSlices);
    cell(b.avgSlices);
    cell(b.minSlices);
    cell(b.maxSlices);
    cell(b.stdSlices);
    cell(b.noPaddedSlices);
    cell(b.noUnpaddedSlices);
    cell(b.noSlices * b.sliceSize);
    cell(b.noUnpaddedSlices * b.sliceSize);
    cell(b.compressionRatio);
    return o.str();
}

}  // namespace blest
