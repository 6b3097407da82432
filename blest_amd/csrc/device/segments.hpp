// Contention-free frontier lists and counters.
//
// Same-address atomics from many workgroups serialise at the device's coherence point
// (~80 ns each on MI355X), so producers never share a counter.  The bitmap words are cut
// into `segs` contiguous segments, one per producer workgroup.  Because the vsets of an
// rset are numbered contiguously and rsets ascend with the words, the vsets a segment
// can emit lie in the fixed list range [segBase[s], segBase[s+1]); the producer writes
// its vsets compactly from segBase[s] (block-wide scan, no atomics) and stores the
// count in segCount[s].  Per-segment vertex and edge counts are stored the same way.
// Consumers read the segs counts and take a prefix in LDS.
#pragma once

#include "dev_common.hpp"

namespace blest {
namespace dev {

constexpr int kMaxSegs = 2048;
constexpr int kProducerBlock = 512;

// A level kernel runs one workgroup per segment, and its level ends with its slowest
// workgroup, so the default grid (4 workgroups of 512 threads per CU = 8 waves per SIMD)
// must be resident at once: at 7 waves per SIMD (more than 64 VGPRs, or more than ~100
// SGPRs) a CU holds 3 of them and the last quarter of the grid starts only when others
// finish (measured: 27-33 us late on the large levels of a scale-22 BFS).
#define BLEST_RESIDENT_GRID __attribute__((amdgpu_waves_per_eu(8)))

// An active rset with more than kChunkVsets vsets (a hub's: up to thousands of vsets,
// and an ordering such as Gorder puts hubs next to each other, so one segment can hold
// 10^5 vsets of them against ~10^3 on average) is not listed vset by vset but as chunk
// entries (first vset, count <= kChunkVsets) in a second segmented list: its producer
// writes up to 16 times fewer entries (a segment's list is at most 16 per rset), and
// chunks hold >= kChunkVsets / 2 vsets on average.  The BFS producer and consumer use
// chunks; the multi-source and label-propagation lists stay per vset.
constexpr uint32_t kChunkVsets = 16;

struct SegView {
    uint32_t segs;              // producer workgroups
    uint32_t wps;               // bitmap words per segment
    const uint32_t* segBase;    // segs + 1: first list slot of each segment
    const uint32_t* chunkBase;  // segs + 1: first chunk slot of each segment
};

// Per-slot per-segment counters (slot = level % 3).
struct SegCounters {
    uint32_t* count;          // [3][segs] list entries
    uint32_t* news;           // [3][segs] vertices (or vertex-root pairs) discovered
    unsigned long long* edges;  // [3][segs] their degree sum
    uint32_t* chunks;         // [3][segs] chunk entries
};

// Chunk entries an rset with cnt vsets contributes (0 when it is listed per vset).
__host__ __device__ __forceinline__ uint32_t rset_chunks(uint32_t cnt) {
    return cnt > kChunkVsets ? (cnt + kChunkVsets - 1) / kChunkVsets : 0u;
}

// Exclusive scan over the block (blockDim.x <= 1024, multiple of 64); returns the
// thread's exclusive prefix and the block total.  `tmp` holds >= 16 entries.
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned x, unsigned* tmp, unsigned& total) {
    const unsigned lane = lane_id(), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const unsigned incl = wave_inclusive_scan(x);
    __syncthreads();
    if (lane == kWave - 1) tmp[wave] = incl;
    __syncthreads();
    unsigned before = 0, all = 0;
    for (unsigned i = 0; i < nw; ++i) {
        const unsigned t = tmp[i];
        before += (i < wave) ? t : 0u;
        all += t;
    }
    total = all;
    return before + incl - x;
}

// Two exclusive scans over the block through one LDS exchange (the barriers of one
// scan): returns x's exclusive prefix, y's in ey, and the totals.  `tmp` holds >= 32.
__device__ __forceinline__ unsigned block_exclusive_scan2(unsigned x, unsigned y, unsigned* tmp, unsigned& total,
                                                          unsigned& ey, unsigned& ytotal) {
    const unsigned lane = lane_id(), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const unsigned ix = wave_inclusive_scan(x), iy = wave_inclusive_scan(y);
    __syncthreads();
    if (lane == kWave - 1) {
        tmp[wave] = ix;
        tmp[16 + wave] = iy;
    }
    __syncthreads();
    unsigned bx = 0, by = 0, ax = 0, ay = 0;
    for (unsigned i = 0; i < nw; ++i) {
        const unsigned tx = tmp[i], ty = tmp[16 + i];
        bx += (i < wave) ? tx : 0u;
        by += (i < wave) ? ty : 0u;
        ax += tx;
        ay += ty;
    }
    total = ax;
    ytotal = ay;
    ey = by + iy - y;
    return bx + ix - x;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long x, unsigned long long* tmp) {
    const unsigned lane = lane_id(), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    x = wave_sum_u64(x);
    __syncthreads();
    if (lane == 0) tmp[wave] = x;
    __syncthreads();
    unsigned long long s = 0;
    for (unsigned i = 0; i < nw; ++i) s += tmp[i];
    return s;
}

// Loads a [segs] count row into LDS as an inclusive prefix (pre[i] = sum of counts
// 0..i); returns the total.  pre holds kMaxSegs entries.
__device__ __forceinline__ unsigned load_prefix(const uint32_t* row, uint32_t segs, unsigned* pre, unsigned* tmp) {
    unsigned carry = 0;
    for (uint32_t base = 0; base < segs; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const unsigned x = i < segs ? row[i] : 0u;
        unsigned tot;
        const unsigned ex = block_exclusive_scan(x, tmp, tot);
        if (i < segs) pre[i] = carry + ex + x;
        carry += tot;
        __syncthreads();
    }
    __syncthreads();
    return carry;
}

// Segment holding global entry index e (pre = inclusive prefix, e < total).
__device__ __forceinline__ uint32_t find_segment(const unsigned* pre, uint32_t segs, unsigned e) {
    uint32_t lo = 0, hi = segs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] > e) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// List entry with global index e.
__device__ __forceinline__ uint32_t list_entry(const uint32_t* list, const SegView& sv, const unsigned* pre,
                                               unsigned e) {
    const uint32_t s = find_segment(pre, sv.segs, e);
    const unsigned before = s ? pre[s - 1] : 0u;
    return list[sv.segBase[s] + (e - before)];
}

// Chunk entry with global index c (cpre = inclusive prefix of the chunk counts).
__device__ __forceinline__ uint2 chunk_entry(const uint2* chunks, const SegView& sv, const unsigned* cpre,
                                             unsigned c) {
    const uint32_t s = find_segment(cpre, sv.segs, c);
    const unsigned before = s ? cpre[s - 1] : 0u;
    return chunks[sv.chunkBase[s] + (c - before)];
}

// Number of vsets of the rsets with a nonzero byte in word w's bits.
__device__ __forceinline__ unsigned word_vsets(const uint32_t* rsetUnit, uint32_t w, unsigned long long bits) {
    unsigned cnt = 0;
    if (bits)
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if ((bits >> (8 * b)) & 0xFFull) cnt += rsetUnit[w * 8 + b + 1] - rsetUnit[w * 8 + b];
    return cnt;
}

// Writes, at out, the vsets of the rsets active in words [first, last) (in word and
// rset order; bitsOf(w) gives word w's frontier bits) and returns their number.  Lane
// = one of the 64 rsets of 8 consecutive words; each step's entries are then written
// by all 64 lanes (owner search), so an rset with thousands of vsets (a hub's) costs
// the wave as many stores as any other 64 entries.  The whole wave must call.
template <class BitsOf>
__device__ __forceinline__ unsigned wave_write_vsets(const uint32_t* rsetUnit, uint32_t first, uint32_t last,
                                                     BitsOf bitsOf, uint32_t* out) {
    const unsigned lane = lane_id();
    unsigned written = 0;
    for (uint32_t c0 = first; c0 < last; c0 += kWave / 8) {
        const uint32_t w = c0 + lane / 8, b = lane % 8;
        uint32_t start = 0, cnt = 0;
        const unsigned long long bits = bitsOf(min(w, last - 1));  // every lane calls (bitsOf may shuffle)
        if (w < last) {
            if ((bits >> (8 * b)) & 0xFFull) {
                start = rsetUnit[w * 8 + b];
                cnt = rsetUnit[w * 8 + b + 1] - start;
            }
        }
        const unsigned incl = wave_inclusive_scan(cnt);
        const unsigned excl = incl - cnt;
        const unsigned total = __shfl(incl, kWave - 1, kWave);
        for (unsigned t0 = 0; t0 < total; t0 += kWave) {
            const unsigned t = t0 + lane;
            const unsigned owner = wave_owner(excl, min(t, total - 1));
            const uint32_t os = __shfl(start, owner, kWave);
            const unsigned oe = __shfl(excl, owner, kWave);
            if (t < total) out[written + t] = os + (t - oe);
        }
        written += total;
    }
    return written;
}

// Block-wide form for a producer workgroup's whole segment [ws, we): thread = one rset
// of blockDim/8 consecutive words per chunk; after one block scan each rset of fewer than
// kRunVsets vsets writes its own entries (a handful on R-MAT graphs), and the larger ones
// (a hub's: up to thousands; Gorder puts hubs side by side, so one segment can hold 10^5
// vsets of them) are written after them as runs, start, start + 1, ..., dealt out to the
// waves, each run written by one wave's lanes.  (An owner search over the chunk's prefix
// in LDS, every thread writing entries a workgroup apart, cost a galloping search of
// dependent LDS reads per entry: a 128-root update over a segment of 1.7e5 hub vsets took
// 130-150 us.)  prep(c0) runs on every thread before a chunk's bits are read (e.g. to
// stage them in LDS).  tmp holds 32 entries.  Returns the segment's entry count.
constexpr uint32_t kRunVsets = 32;

template <class Prep, class BitsOf>
__device__ __forceinline__ unsigned block_write_vsets(const uint32_t* rsetUnit, uint32_t ws, uint32_t we, Prep prep,
                                                      BitsOf bitsOf, uint32_t* out, unsigned* tmp) {
    __shared__ uint32_t sRunStart[kProducerBlock], sRunCnt[kProducerBlock], sRunOff[kProducerBlock];
    __shared__ unsigned sRuns;
    const unsigned t = threadIdx.x, nb = blockDim.x;
    unsigned written = 0;
    for (uint32_t c0 = ws; c0 < we; c0 += nb / 8) {
        prep(c0);
        const uint32_t w = c0 + t / 8, b = t % 8;
        uint32_t start = 0, cnt = 0;
        if (w < we) {
            const unsigned long long bits = bitsOf(w);
            if ((bits >> (8 * b)) & 0xFFull) {
                start = rsetUnit[w * 8 + b];
                cnt = rsetUnit[w * 8 + b + 1] - start;
            }
        }
        if (t == 0) sRuns = 0;  // ordered before the claims below by the scan's barriers
        const bool run = cnt >= kRunVsets;
        unsigned total, rex, rtotal;
        const unsigned ex = block_exclusive_scan2(run ? 0u : cnt, run ? cnt : 0u, tmp, total, rex, rtotal);
        if (run) {
            const unsigned k = atomicAdd(&sRuns, 1u);  // queue order is free: offsets come from the scan
            sRunStart[k] = start;
            sRunCnt[k] = cnt;
            sRunOff[k] = total + rex;
        } else {
            for (uint32_t j = 0; j < cnt; ++j) out[written + ex + j] = start + j;
        }
        __syncthreads();
        const unsigned nr = sRuns, wave = t / kWave, nwv = nb / kWave, lane = t % kWave;
        for (unsigned k = wave; k < nr; k += nwv) {
            const uint32_t s0 = sRunStart[k], c = sRunCnt[k];
            uint32_t* o2 = out + written + sRunOff[k];
            for (uint32_t j = lane; j < c; j += kWave) o2[j] = s0 + j;
        }
        written += total + rtotal;
        __syncthreads();  // the next chunk reuses the run queue
    }
    return written;
}

// Word range [first, last) of wave `wib` when a segment [ws, we) is cut into
// contiguous runs, one per wave of the workgroup.
__device__ __forceinline__ void wave_words(uint32_t ws, uint32_t we, unsigned wib, unsigned wpb, uint32_t& first,
                                           uint32_t& last) {
    const uint32_t per = (we - ws + wpb - 1) / wpb;
    first = min(we, ws + wib * per);
    last = min(we, first + per);
}

}  // namespace dev
}  // namespace blest
