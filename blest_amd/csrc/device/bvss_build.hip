// BVSS construction on the GPU (slice size 8, 4 masks): the same five arrays, bit for
// bit, as the host construction (bvss.cpp, the reference's BVSS::constructFromCSCMatrix,
// BVSS.cuh:185-467), built by sorting instead of a per-rset k-way merge.
//
//   1. keys: every stored entry (row i, column j) becomes key (i << 3) | (j & 7); the
//      entries of rset r (columns 8r..8r+7) are contiguous in CSC order, so a segmented
//      radix sort with one segment per rset orders each rset's entries by row, the
//      column bits of one row next to each other.
//   2. slices: a slice starts at a segment head or where the row changes; a scan of
//      those flags numbers the slices, the head of each slice ORs its entries' bits
//      into the 8-bit mask.  Rsets are taken in batches of at most 2^28 entries.
//   3. layout: per rset the vset count and slot count of the reference's rule
//      (complete vsets of 32 lanes x 4 positions; the leftover padded to 128 slots
//      with FULL_PADDING, else floor(left / 4) lanes strided by that count plus a last
//      lane padded to 4), prefix sums for realPtrs and the slot offsets, then every
//      real slice scattered to its (lane, position) slot.  Byte p of mask word t is
//      the mask of slot 4t + p, so a slice's mask byte index equals its slot index.
//   4. the update divergence (BVSS.cuh:420-467): each vset's term in FP64 with the host
//      loop's operation order and no fused operations, summed over vsets on the host in
//      vset order, as the host construction does.
#include <hipcub/hipcub.hpp>

#include <memory>
#include <vector>

#include "../bvss.hpp"
#include "dgraph.hpp"

namespace blest {
namespace dev {

constexpr uint32_t kBbLanes = 32, kBbMasks = 4, kBbFull = kBbLanes * kBbMasks;  // 128 slots per vset

__global__ void bb_keys_kernel(const uint32_t* colPtrs, const uint32_t* rows, uint32_t c0, uint32_t c1, uint64_t e0,
                               uint32_t* keys) {
    const unsigned lane = lane_id();
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t c = c0 + wave; c < c1; c += nwaves) {
        const uint32_t b = colPtrs[c], e = colPtrs[c + 1];
        for (uint32_t p = b + lane; p < e; p += kWave) keys[p - e0] = (rows[p] << 3) | (c & 7u);
    }
}

// off[i] = first entry of rset r0 + i relative to the batch (i = 0..nseg)
__global__ void bb_offsets_kernel(const uint32_t* colPtrs, uint32_t n, uint32_t r0, uint32_t nseg, uint64_t e0,
                                  uint32_t* off) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= nseg; i += gridDim.x * blockDim.x) {
        const uint64_t c = min(uint64_t(r0 + i) * 8, uint64_t(n));
        off[i] = uint32_t(colPtrs[c] - e0);
    }
}

__global__ void bb_fill_u8_kernel(uint8_t* a, uint8_t v, size_t count) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        a[i] = v;
}

__global__ void bb_fill_u32_kernel(uint32_t* a, uint32_t v, size_t count) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        a[i] = v;
}

__global__ void bb_heads_kernel(const uint32_t* off, uint32_t nseg, uint8_t* head) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nseg; i += gridDim.x * blockDim.x)
        if (off[i] < off[i + 1]) head[off[i]] = 1;
}

// flag[e] = 1 where a slice starts (a segment head, or a row other than the previous entry's)
__global__ void bb_flags_kernel(const uint32_t* keys, const uint8_t* head, uint32_t items, uint32_t* flag) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < items; e += gridDim.x * blockDim.x)
        flag[e] = (e == 0 || head[e] || (keys[e] >> 3) != (keys[e - 1] >> 3)) ? 1u : 0u;
}

// The head entry of each slice writes its row and the OR of its entries' column bits (at
// most 8 entries, contiguous up to the next flag).
__global__ void bb_slices_kernel(const uint32_t* keys, const uint32_t* flag, const uint32_t* pos, uint32_t items,
                                 uint64_t sbase, uint32_t* sRows, uint8_t* sMasks) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < items; e += gridDim.x * blockDim.x) {
        if (!flag[e]) continue;
        uint32_t m = 1u << (keys[e] & 7u);
        for (uint32_t j = e + 1; j < items && !flag[j]; ++j) m |= 1u << (keys[j] & 7u);
        sRows[sbase + pos[e]] = keys[e] >> 3;
        sMasks[sbase + pos[e]] = uint8_t(m);
    }
}

// slices of rset r0 + i: slice number of its first entry to that of the next rset's
__global__ void bb_counts_kernel(const uint32_t* off, uint32_t nseg, const uint32_t* pos, uint32_t items,
                                 uint32_t total, uint32_t* cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nseg; i += gridDim.x * blockDim.x) {
        const uint32_t a = off[i] < items ? pos[off[i]] : total;
        const uint32_t b = off[i + 1] < items ? pos[off[i + 1]] : total;
        cnt[i] = b - a;
    }
}

__device__ __forceinline__ uint64_t bb_vset_slots(uint64_t c, bool fullPadding, uint64_t vi) {
    const uint64_t complete = c / kBbFull;
    if (vi < complete || fullPadding) return kBbFull;
    const uint64_t left = c - complete * kBbFull, stride = left / kBbMasks;
    return stride * kBbMasks + ((left - stride * kBbMasks) ? kBbMasks : 0);
}

// vsets and slots of each rset (trailing entry R = 0, so exclusive scans end in the totals)
__global__ void bb_shape_kernel(const uint32_t* cnt, uint32_t R, int fullPadding, uint32_t* vsets,
                                unsigned long long* slots, unsigned long long* cnt64) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= R; r += gridDim.x * blockDim.x) {
        const uint64_t c = r < R ? cnt[r] : 0;
        const uint64_t complete = c / kBbFull, left = c - complete * kBbFull;
        vsets[r] = uint32_t(complete + (left ? 1 : 0));
        slots[r] = complete * kBbFull + (left ? bb_vset_slots(c, fullPadding != 0, complete) : 0);
        cnt64[r] = c;
    }
}

__global__ void bb_vsets_kernel(const uint32_t* cnt, const uint32_t* realPtrs, const unsigned long long* rsetSlot,
                                uint32_t R, int fullPadding, unsigned long long* sliceSetPtrs, uint32_t* v2r) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x) {
        uint64_t slot = rsetSlot[r];
        for (uint32_t v = realPtrs[r]; v < realPtrs[r + 1]; ++v) {
            slot += bb_vset_slots(cnt[r], fullPadding != 0, v - realPtrs[r]);
            sliceSetPtrs[v + 1] = slot;
            v2r[v] = r;
        }
    }
}

// Slot (relative to the rset's first slot) of the rset's i-th slice out of c.
__device__ __forceinline__ uint64_t bb_slot_of(uint64_t i, uint64_t c, bool fullPadding) {
    const uint64_t complete = c / kBbFull, ls = complete * kBbFull;
    if (i < ls) {
        const uint64_t vi = i / kBbFull, j = i % kBbFull;
        return vi * kBbFull + (j % kBbLanes) * kBbMasks + j / kBbLanes;  // lane j % 32, position j / 32
    }
    const uint64_t j = i - ls;
    if (fullPadding) return ls + (j % kBbLanes) * kBbMasks + j / kBbLanes;
    const uint64_t stride = (c - ls) / kBbMasks;
    if (j < stride * kBbMasks) return ls + (j % stride) * kBbMasks + j / stride;  // lane j % stride, position j / stride
    return ls + stride * kBbMasks + (j - stride * kBbMasks);                       // the last lane, padded to 4
}

__global__ void bb_scatter_kernel(const uint32_t* sRows, const uint8_t* sMasks, unsigned long long sreal,
                                  const unsigned long long* sliceStart, const unsigned long long* rsetSlot, uint32_t R,
                                  int fullPadding, uint32_t* rowIds, uint8_t* maskBytes) {
    for (uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; s < sreal; s += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t lo = 0, hi = R;  // rset r with sliceStart[r] <= s < sliceStart[r + 1]
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sliceStart[mid] <= s) lo = mid;
            else hi = mid;
        }
        const uint64_t c = sliceStart[lo + 1] - sliceStart[lo];
        const uint64_t slot = rsetSlot[lo] + bb_slot_of(s - sliceStart[lo], c, fullPadding != 0);
        rowIds[slot] = sRows[s];
        maskBytes[slot] = sMasks[s];
    }
}

// The vset's term of computeUpdateDivergence in the host loop's order, every operation
// rounded on its own (no fused multiply-add), so that the FP64 value matches the host's.
__global__ void bb_divergence_kernel(const unsigned long long* sliceSetPtrs, const uint32_t* rowIds, uint32_t V,
                                     double* term) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint64_t base = sliceSetPtrs[v], size = sliceSetPtrs[v + 1] - base;
        double vmean = 0.0;
        for (uint32_t p = 0; p < kBbMasks; ++p) {
            double mean = 0.0, sd = 0.0;
            uint32_t k = 0;
            for (uint32_t t = 0; t < kBbLanes; ++t) {
                const uint64_t sl = uint64_t(t) * kBbMasks + p;
                if (sl < size && rowIds[base + sl] != 0xFFFFFFFFu) {
                    mean = __dadd_rn(mean, double(rowIds[base + sl]));
                    ++k;
                }
            }
            if (k) mean = __ddiv_rn(mean, double(k));
            for (uint32_t t = 0; t < kBbLanes; ++t) {
                const uint64_t sl = uint64_t(t) * kBbMasks + p;
                if (sl < size && rowIds[base + sl] != 0xFFFFFFFFu) {
                    const double d = __dadd_rn(double(rowIds[base + sl]), -mean);
                    sd = __dadd_rn(sd, __dmul_rn(d, d));
                }
            }
            if (k) sd = __ddiv_rn(sd, double(k));
            vmean = __dadd_rn(vmean, __dsqrt_rn(sd));
        }
        term[v] = __ddiv_rn(vmean, double(kBbMasks));
    }
}

}  // namespace dev

namespace {

// Device buffers of one construction, freed together.
struct Scratch {
    std::vector<void*> ptrs;
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        BLEST_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

constexpr uint64_t kBatchItems = uint64_t(1) << 28;

dim3 grid_for(uint64_t items, unsigned block = 256) {
    return dim3(unsigned(std::max<uint64_t>(1, std::min<uint64_t>((items + block - 1) / block, 65536))));
}

template <class T>
T read_one(const T* d, hipStream_t s) {
    T v{};
    BLEST_HIP(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s));
    BLEST_HIP(hipStreamSynchronize(s));
    return v;
}

BVSS bvss_build_on_device(const CSC& g, bool fullPadding, int device) {
    BLEST_HIP(hipSetDevice(device));
    StreamGuard sg;
    BLEST_HIP(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    hipStream_t st = sg.s;
    Scratch sc;
    const uint32_t n = g.n;
    const uint64_t nnz = g.nnz();
    if (uint64_t(n) > (uint64_t(1) << 29)) throw Error(BLEST_ERR_UNSUPPORTED, "device BVSS build needs n <= 2^29");
    const uint32_t R = uint32_t((uint64_t(n) + 7) / 8);
    uint32_t* dColPtrs = sc.alloc<uint32_t>(size_t(n) + 1);
    uint32_t* dRows = sc.alloc<uint32_t>(nnz);
    BLEST_HIP(hipMemcpyAsync(dColPtrs, g.colPtrs.data(), 4 * (size_t(n) + 1), hipMemcpyHostToDevice, st));
    if (nnz) BLEST_HIP(hipMemcpyAsync(dRows, g.rows.data(), 4 * nnz, hipMemcpyHostToDevice, st));

    // rset batches of at most kBatchItems entries (one rset may exceed it alone)
    std::vector<std::pair<uint32_t, uint32_t>> batches;
    for (uint32_t r0 = 0; r0 < R;) {
        uint32_t r1 = r0 + 1;
        const uint64_t e0 = g.colPtrs[size_t(r0) * 8];
        while (r1 < R && g.colPtrs[std::min<uint64_t>(uint64_t(r1 + 1) * 8, n)] - e0 <= kBatchItems) ++r1;
        batches.emplace_back(r0, r1);
        r0 = r1;
    }
    uint64_t maxItems = 1, maxSegs = 1;
    for (auto [r0, r1] : batches) {
        maxItems = std::max<uint64_t>(maxItems, g.colPtrs[std::min<uint64_t>(uint64_t(r1) * 8, n)] -
                                                    g.colPtrs[size_t(r0) * 8]);
        maxSegs = std::max<uint64_t>(maxSegs, r1 - r0);
    }
    if (maxItems > uint64_t(0x7FFFFFFF)) throw Error(BLEST_ERR_UNSUPPORTED, "an rset holds more than 2^31 entries");
    uint32_t* keysIn = sc.alloc<uint32_t>(maxItems);
    uint32_t* keysOut = sc.alloc<uint32_t>(maxItems);
    uint32_t* flag = sc.alloc<uint32_t>(maxItems);
    uint32_t* pos = sc.alloc<uint32_t>(maxItems);
    uint8_t* head = sc.alloc<uint8_t>(maxItems);
    uint32_t* off = sc.alloc<uint32_t>(maxSegs + 1);
    uint32_t* cnt = sc.alloc<uint32_t>(size_t(R) + 1);
    uint32_t* sRows = sc.alloc<uint32_t>(nnz);
    uint8_t* sMasks = sc.alloc<uint8_t>(nnz);
    int endBit = 3;
    while ((uint64_t(1) << (endBit - 3)) < n) ++endBit;
    size_t sortBytes = 0, scanBytes = 0;
    BLEST_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, sortBytes, keysIn, keysOut, int(maxItems),
                                                         int(maxSegs), off, off + 1, 0, endBit, st));
    BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, flag, pos, int(maxItems), st));
    void* temp = sc.alloc<uint8_t>(std::max(sortBytes, scanBytes));

    uint64_t sreal = 0;
    for (auto [r0, r1] : batches) {
        const uint32_t nseg = r1 - r0;
        const uint32_t c0 = r0 * 8, c1 = uint32_t(std::min<uint64_t>(uint64_t(r1) * 8, n));
        const uint64_t e0 = g.colPtrs[c0];
        const uint32_t items = uint32_t(g.colPtrs[c1] - e0);
        hipLaunchKernelGGL(dev::bb_offsets_kernel, grid_for(nseg + 1), dim3(256), 0, st, dColPtrs, n, r0, nseg, e0,
                           off);
        if (items == 0) {
            hipLaunchKernelGGL(dev::bb_fill_u32_kernel, grid_for(nseg), dim3(256), 0, st, cnt + r0, 0u, size_t(nseg));
            BLEST_HIP(hipGetLastError());
            continue;
        }
        hipLaunchKernelGGL(dev::bb_keys_kernel, grid_for(uint64_t(c1 - c0) * 64), dim3(256), 0, st, dColPtrs, dRows,
                           c0, c1, e0, keysIn);
        hipLaunchKernelGGL(dev::bb_fill_u8_kernel, grid_for(items), dim3(256), 0, st, head, uint8_t(0), size_t(items));
        hipLaunchKernelGGL(dev::bb_heads_kernel, grid_for(nseg), dim3(256), 0, st, off, nseg, head);
        BLEST_HIP(hipGetLastError());
        size_t tb = sortBytes;
        BLEST_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(temp, tb, keysIn, keysOut, int(items), int(nseg), off,
                                                             off + 1, 0, endBit, st));
        hipLaunchKernelGGL(dev::bb_flags_kernel, grid_for(items), dim3(256), 0, st, keysOut, head, items, flag);
        BLEST_HIP(hipGetLastError());
        tb = scanBytes;
        BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(temp, tb, flag, pos, int(items), st));
        const uint32_t total = read_one(pos + items - 1, st) + read_one(flag + items - 1, st);
        hipLaunchKernelGGL(dev::bb_slices_kernel, grid_for(items), dim3(256), 0, st, keysOut, flag, pos, items, sreal,
                           sRows, sMasks);
        hipLaunchKernelGGL(dev::bb_counts_kernel, grid_for(nseg), dim3(256), 0, st, off, nseg, pos, items, total,
                           cnt + r0);
        BLEST_HIP(hipGetLastError());
        sreal += total;
    }

    // layout
    uint32_t* vsets = sc.alloc<uint32_t>(size_t(R) + 1);
    unsigned long long* slots = sc.alloc<unsigned long long>(size_t(R) + 1);
    unsigned long long* cnt64 = sc.alloc<unsigned long long>(size_t(R) + 1);
    uint32_t* realPtrs = sc.alloc<uint32_t>(size_t(R) + 1);
    unsigned long long* rsetSlot = sc.alloc<unsigned long long>(size_t(R) + 1);
    unsigned long long* sliceStart = sc.alloc<unsigned long long>(size_t(R) + 1);
    hipLaunchKernelGGL(dev::bb_shape_kernel, grid_for(uint64_t(R) + 1), dim3(256), 0, st, cnt, R, int(fullPadding),
                       vsets, slots, cnt64);
    BLEST_HIP(hipGetLastError());
    size_t s1 = 0, s2 = 0;
    BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, s1, vsets, realPtrs, int(R) + 1, st));
    BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, s2, slots, rsetSlot, int(R) + 1, st));
    void* temp2 = sc.alloc<uint8_t>(std::max(s1, s2));
    BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(temp2, s1, vsets, realPtrs, int(R) + 1, st));
    s2 = std::max(s1, s2);
    BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(temp2, s2, slots, rsetSlot, int(R) + 1, st));
    s2 = std::max(s1, s2);
    BLEST_HIP(hipcub::DeviceScan::ExclusiveSum(temp2, s2, cnt64, sliceStart, int(R) + 1, st));
    const uint32_t V = read_one(realPtrs + R, st);
    const uint64_t S = read_one(rsetSlot + R, st);
    if (read_one(sliceStart + R, st) != sreal) throw Error(BLEST_ERR_DEVICE, "device BVSS build: slice count mismatch");
    unsigned long long* dSsp = sc.alloc<unsigned long long>(size_t(V) + 1);
    uint32_t* dV2r = sc.alloc<uint32_t>(V);
    uint32_t* dRowIds = sc.alloc<uint32_t>(S);
    uint32_t* dMasks = sc.alloc<uint32_t>(S / 4);
    double* dTerm = sc.alloc<double>(V);
    hipLaunchKernelGGL(dev::bb_fill_u32_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(dSsp), 0u,
                       size_t(2));  // sliceSetPtrs[0] = 0 (64-bit zero)
    hipLaunchKernelGGL(dev::bb_vsets_kernel, grid_for(R), dim3(256), 0, st, cnt, realPtrs, rsetSlot, R,
                       int(fullPadding), dSsp, dV2r);
    hipLaunchKernelGGL(dev::bb_fill_u32_kernel, grid_for(S), dim3(256), 0, st, dRowIds, 0xFFFFFFFFu, size_t(S));
    hipLaunchKernelGGL(dev::bb_fill_u32_kernel, grid_for(S / 4), dim3(256), 0, st, dMasks, 0u, size_t(S / 4));
    hipLaunchKernelGGL(dev::bb_scatter_kernel, grid_for(sreal), dim3(256), 0, st, sRows, sMasks,
                       (unsigned long long)sreal, sliceStart, rsetSlot, R, int(fullPadding), dRowIds,
                       reinterpret_cast<uint8_t*>(dMasks));
    hipLaunchKernelGGL(dev::bb_divergence_kernel, grid_for(V), dim3(256), 0, st, dSsp, dRowIds, V, dTerm);
    BLEST_HIP(hipGetLastError());

    BVSS b;
    b.n = n;
    b.sliceSize = 8;
    b.noMasks = 4;
    b.fullPadding = fullPadding;
    b.noRealSliceSets = R;
    b.noVirtualSliceSets = V;
    b.noSlices = S;
    b.sliceSetPtrs.resize(size_t(V) + 1);
    b.virtualToReal.resize(V);
    b.realPtrs.resize(size_t(R) + 1);
    b.rowIds.resize(S);
    b.masks.resize(S / 4);
    std::vector<double> term(V);
    BLEST_HIP(hipMemcpyAsync(b.sliceSetPtrs.data(), dSsp, 8 * (size_t(V) + 1), hipMemcpyDeviceToHost, st));
    if (V) BLEST_HIP(hipMemcpyAsync(b.virtualToReal.data(), dV2r, 4 * size_t(V), hipMemcpyDeviceToHost, st));
    BLEST_HIP(hipMemcpyAsync(b.realPtrs.data(), realPtrs, 4 * (size_t(R) + 1), hipMemcpyDeviceToHost, st));
    if (S) {
        BLEST_HIP(hipMemcpyAsync(b.rowIds.data(), dRowIds, 4 * size_t(S), hipMemcpyDeviceToHost, st));
        BLEST_HIP(hipMemcpyAsync(b.masks.data(), dMasks, size_t(S), hipMemcpyDeviceToHost, st));
    }
    if (V) BLEST_HIP(hipMemcpyAsync(term.data(), dTerm, 8 * size_t(V), hipMemcpyDeviceToHost, st));
    BLEST_HIP(hipStreamSynchronize(st));
    double div = 0;
    for (uint32_t v = 0; v < V; ++v) div += term[v];  // the host construction's summation order
    b.updateDivergence = V ? div / V : 0.0;
    bvss_compute_stats(b);
    return b;
}

}  // namespace
}  // namespace blest

using namespace blest;

extern "C" blest_status blest_bvss_build_device(const blest_csc* gh, uint32_t slice_size, uint32_t no_masks,
                                                int full_padding, int device, blest_bvss** out) {
    return guarded([&] {
        if (!gh || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        bvss_check_shape(slice_size, no_masks);
        if (slice_size != 8 || no_masks != 4)
            throw Error(BLEST_ERR_UNSUPPORTED, "the device BVSS build covers slice size 8 with 4 masks");
        const CSC& g = csc_of(gh);
        const bool fp = full_padding < 0 ? !g.isSocial : full_padding != 0;
        *out = bvss_wrap(bvss_build_on_device(g, fp, device));
    });
}
