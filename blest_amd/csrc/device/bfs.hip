// Single-source direction-optimising BFS over the HBM-resident BVSS (gfx950, wave64).
//
// Reference path: BVSSBFSKernel::hostCode and its persistent cooperative kernels
// (BFS/BVSSBFSKernel.cuh:33-284 top-down/bottom-up switching, 531-691 top-down only,
// 1283-1608 host side).  On MI355X a grid-wide barrier costs 7-33 us while a kernel
// boundary on one stream costs ~2.3 us (1.6 us replayed from a graph), so the level
// loop is a stream of small kernels instead of one cooperative launch: two kernels
// per level, every decision taken on the device, the host only polling a pinned
// `done` flag once per chunk of levels.  Frontier lists and counters are segmented
// per producer workgroup (segments.hpp) so that no level funnels through one atomic.
//
// Level L (frontier F = vertices at distance L-1):
//   expand  every block derives the same direction from the previous level's counts
//           (one batch of loads, level_common.hpp);
//           top-down: each wave step takes m list entries (vsets) or chunk slots (a
//             hub rset's vsets, 16 to a chunk), scans their quad counts across the
//             wave and streams the union of their quads (every lane busy, loads
//             coalesced); a slice hits when (mask & frontier byte) != 0, and a hit row
//             not yet visited gets a plain byte store in `nextB` (no atomics);
//           bottom-up: workgroup per segment; a lane owns one vertex in each of 4 words
//             of its wave and scans their in-neighbours one vertex after another until
//             one is in the frontier (1, 2, then 4 entries per step); a scan open after
//             32 entries is finished 64 entries at a time by a whole wave.
//   update  workgroup per segment, thread per rset: top-down packs each word's 64
//           discovery bytes into the new frontier word and folds it into visited /
//           levels; both directions list the vsets of the rsets with new vertices in
//           the segment's list range (hub rsets as chunks in its chunk range).
// Direction: top-down -> bottom-up when the frontier's edges exceed the unexplored
// edges / alpha; back when the frontier shrinks below n / beta.  Levels are unique
// BFS distances, so every schedule produces the same result.
#include <cstring>

#include "dgraph.hpp"
#include "level_common.hpp"

namespace blest {
namespace dev {

struct BfsArgs {
    int32_t* levels;
    unsigned long long* visited;
    unsigned long long* front0;
    unsigned long long* front1;
    uint8_t* nextB;  // top-down discoveries, one byte per vertex (W * 64 bytes)
    uint32_t* list0;
    uint32_t* list1;
    uint2* chunk0;  // chunk lists (rsets of more than kChunkVsets vsets)
    uint2* chunk1;
    Ctrl* ctrl;
    SegView sv;
    SegCounters sc;
    uint32_t src;
    int direction;
    double alpha, beta;
    int levelFromCtrl;  // 1: the kernel's L argument is relative to ctrl->levelBase
    unsigned fuseQuads;  // quads a fused small level may stream (0: levels never fused)
};

// a[k] for a per-lane k (a select chain: the array stays in registers); 0 for k >= N.
template <int N>
__device__ __forceinline__ uint32_t pick(const uint32_t (&a)[N], int k) {
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) r = (i == k) ? a[i] : r;
    return r;
}

__device__ __forceinline__ uint32_t front_byte(const unsigned long long* f, uint32_t r) {
    return reinterpret_cast<const uint8_t*>(f)[r];
}

// Slice (row, hit) of a frontier column: a row not yet visited is marked with a plain
// byte store (idempotent, so no atomic is needed).  A 64-bit atomicOr into a bitmap word
// per hit serialised at the coherence point when a hub's dense rows met in the same
// words: a hub root's first big top-down level took 46 us with the atomics and 17 us
// without them.  The update kernel packs the bytes into bitmap words.
__device__ __forceinline__ void expand_slice(uint32_t row, uint32_t hit, const unsigned long long* visited,
                                             uint8_t* nextB) {
    if (!hit) return;
    if ((visited[row >> 6] >> (row & 63)) & 1ull) return;
    nextB[row] = 1;
}

__global__ void __launch_bounds__(kProducerBlock) bfs_init_kernel(GraphView g, BfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    const uint32_t sw = a.src >> 6;
    const unsigned long long sb = 1ull << (a.src & 63);
    for (uint32_t w = tid; w < g.W; w += nthreads) {
        const unsigned long long x = (w == sw) ? sb : 0ull;
        a.visited[w] = x;
        a.front1[w] = x;
    }
    unsigned long long* nb64 = reinterpret_cast<unsigned long long*>(a.nextB);
    for (size_t i = tid; i < size_t(g.W) * 8; i += nthreads) nb64[i] = 0ull;
    const uint32_t srcSeg = sw / a.sv.wps;
    for (uint32_t s = tid; s < a.sv.segs; s += nthreads) {
        uint32_t cnt = 0, news = 0, chunks = 0;
        unsigned long long edges = 0;
        if (s == srcSeg) {
            const uint32_t r = a.src >> 3;
            const uint32_t u0 = g.rsetUnit[r], u1 = g.rsetUnit[r + 1];
            chunks = rset_chunks(u1 - u0);
            if (chunks) {
                for (uint32_t j = 0; j < chunks; ++j)
                    a.chunk1[a.sv.chunkBase[s] + j] =
                        make_uint2(u0 + j * kChunkVsets, min(kChunkVsets, u1 - u0 - j * kChunkVsets));
            } else {
                for (uint32_t u = u0; u < u1; ++u) a.list1[a.sv.segBase[s] + (u - u0)] = u;
                cnt = u1 - u0;
            }
            news = 1;
            edges = g.csrPtr ? (g.csrPtr[a.src + 1] - g.csrPtr[a.src]) : 0ull;
        }
        a.sc.count[a.sv.segs + s] = cnt;
        a.sc.chunks[a.sv.segs + s] = chunks;
        a.sc.news[a.sv.segs + s] = news;
        a.sc.edges[a.sv.segs + s] = edges;
    }
    if (tid == 0) {
        a.levels[a.src] = 0;
        level_state_reset(a.ctrl);
    }
}

// ------------------------------------------------------------ fused small levels
// A top-down level whose lists hold a few vsets (a root's first levels, the last levels of
// a traversal, the first levels of a deep one) costs 10-25 us in the two-kernel form: two
// kernel boundaries, a pass over every bitmap word and a few dependent memory round trips
// in each kernel, for a few hundred slices.  bfs_fused_kernel, one workgroup, runs such
// levels back to back with no kernel boundary between them.  It is enqueued after the init
// kernel and after each chunk of level kernels; kernel argument Lk is the enqueued level
// whose inputs it reads (lists, counters, frontier bitmap, state slot), and when it stops
// it leaves the inputs of that same enqueued level, Lk's state carrying the levels it ran
// as its shift, so the expand kernel enqueued as Lk goes on from there.  Per level:
//   queue   the frontier's rsets (the first level's from its lists: the first vset of each
//           listed rset, or the root's rset; later ones from the discoveries), their quad
//           and vset ranges and frontier bytes in LDS.  A level runs here while its queue
//           holds at most kFuseTile rsets and fuseQuads quads (an rset's quads are
//           contiguous);
//   clear   the frontier words of the queue's rsets (their bytes are in LDS now), so the
//           same bitmap receives the level's discoveries;
//   stream  the queue's quads, kFuseQpl per thread per pass, the owner entry found by
//           binary search in the LDS prefix of the entries' quad counts; a slice hit on a
//           row not yet visited claims the row with a 32-bit atomicOr on `visited`; the
//           winner writes the level, sets the row's bit in the bitmap, and when that
//           atomic's old value shows no bit of the row's rset yet, appends the rset.
// A traversal that ends here raises the done flag as level_begin does.  Otherwise, when
// the next level is too big, bottom-up by the direction rule, or the schedule is bottom-up
// after level 1, the workgroup writes that level's lists (in the update kernel's form,
// from the start of the list arrays: one flat range that segment 0's counters describe);
// its frontier is already in the bitmap the expand kernel reads.  Bits of rsets without
// vsets may stay in that bitmap (no list names them, so no queue clears them): their
// vertices have no out-edges, so no direction ever tests them.
constexpr int kFusedBlock = 1024;
constexpr unsigned kFuseTile = 2 * kFusedBlock;  // rsets a fused level's queue may hold
constexpr int kFuseQpl = 4;                      // quads per thread per pass
// While the frontier shrinks (the last levels of a traversal) a level discovers few rows,
// so levels of up to kFuseShrinkScale x fuseQuads quads still run here.
constexpr unsigned kFuseShrinkScale = 4;
constexpr uint32_t kNoRset = 0xFFFFFFFFu;
typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));

// A queue of rsets: entries below kFuseTile in LDS (the level prefix arrays, free once the
// first level's lists are read), later ones in global memory at the same index (the
// top-down discovery bytes `nextB`, all zero between levels: two queues of up to R
// entries fit in its 64 W bytes; an entry used there is zeroed again after use).  Only
// the queue of a level too big to run here reaches past kFuseTile.
struct FusedQueue {
    uint32_t* lds;
    uint32_t* glb;
    __device__ uint32_t get(unsigned i) const { return i < kFuseTile ? lds[i] : glb[i]; }
    __device__ void put(unsigned i, uint32_t r) const {
        if (i < kFuseTile)
            lds[i] = r;
        else
            glb[i] = r;
    }
    __device__ void release(unsigned i) const {
        if (i >= kFuseTile) glb[i] = 0u;
    }
};
static_assert(kFuseTile <= unsigned(kMaxSegs), "an LDS queue lives in a level prefix array");

// Per-entry LDS arrays of a queue of at most kFuseTile rsets.
struct FusedInfo {
    uint32_t* pf;   // exclusive prefix of the entries' quad counts
    uint32_t* q0;   // first quad
    uint32_t* fbq;  // frontier byte, replicated to the 4 slots of a quad
    uint32_t* u0;   // first vset
    uint32_t* uc;   // vsets
};

// Listed vsets (<= 32 quads each) with which a level enters the fused kernel: the entering
// level never exceeds fuseQuads, and holds at most one list entry per thread.
__host__ __device__ constexpr unsigned fuse_vsets(unsigned fuseQuads) {
    return fuseQuads / 32 < unsigned(kFusedBlock) ? fuseQuads / 32 : unsigned(kFusedBlock);
}

// The info of queue entries 2t, 2t + 1 (nq <= kFuseTile); returns the queue's quad total.
// Every thread must call it.
__device__ __forceinline__ unsigned fused_queue_info(const GraphView& g, const uint32_t* q, unsigned nq,
                                                     const unsigned long long* front, const FusedInfo& fi,
                                                     unsigned* tmp32) {
    const unsigned t = threadIdx.x;
    uint32_t u0[2], u1[2], fb[2], qs[2], cnt[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const unsigned i = 2 * t + k;
        const bool ok = i < nq;
        const uint32_t r = ok ? q[i] : 0u;
        u0[k] = ok ? g.rsetUnit[r] : 0u;
        u1[k] = ok ? g.rsetUnit[r + 1] : 0u;
        fb[k] = ok ? front_byte(front, r) : 0u;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const bool live = fb[k] && u1[k] > u0[k];
        qs[k] = live ? g.unitQuad[u0[k]] : 0u;
        cnt[k] = live ? g.unitQuad[u1[k]] - qs[k] : 0u;
    }
    unsigned total;
    const unsigned ex = block_exclusive_scan(cnt[0] + cnt[1], tmp32, total);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const unsigned i = 2 * t + k;
        fi.pf[i] = k ? ex + cnt[0] : ex;
        fi.q0[i] = qs[k];
        fi.fbq[i] = fb[k] * 0x01010101u;
        fi.u0[i] = u0[k];
        fi.uc[i] = u1[k] - u0[k];
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(total);
}

// Lists of queue entries [0, nq) (2 per thread per step), in the order and form the
// update kernel gives them (an rset's vsets, or its chunks beyond kChunkVsets vsets),
// from the start of out / cout; vset ranges from the info when the queue had it
// (nq <= kFuseTile), else loaded.  Every thread must call it.
__device__ __forceinline__ void fused_write_lists(const GraphView& g, const FusedQueue& q, unsigned nq,
                                                  const FusedInfo* fi, uint32_t* out, uint2* cout, unsigned* tmp32,
                                                  unsigned& ltot, unsigned& ctot) {
    const unsigned t = threadIdx.x, nb = blockDim.x;
    ltot = ctot = 0;
    for (unsigned base = 0; base < nq; base += 2 * nb) {
        uint32_t u0[2], cnt[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const unsigned i = base + 2 * t + k;
            if (fi) {
                u0[k] = i < nq ? fi->u0[i] : 0u;
                cnt[k] = i < nq ? fi->uc[i] : 0u;
            } else {
                const uint32_t r = i < nq ? q.get(i) : 0u;
                u0[k] = i < nq ? g.rsetUnit[r] : 0u;
                cnt[k] = i < nq ? g.rsetUnit[r + 1] - u0[k] : 0u;
            }
        }
        const uint32_t ch0 = rset_chunks(cnt[0]), ch1 = rset_chunks(cnt[1]);
        const uint32_t l0 = ch0 ? 0u : cnt[0], l1 = ch1 ? 0u : cnt[1];
        unsigned lsum, csum;
        const unsigned lex = ltot + block_exclusive_scan(l0 + l1, tmp32, lsum);
        const unsigned cex = ctot + block_exclusive_scan(ch0 + ch1, tmp32, csum);
        for (uint32_t j = 0; j < l0; ++j) out[lex + j] = u0[0] + j;
        for (uint32_t j = 0; j < l1; ++j) out[lex + l0 + j] = u0[1] + j;
        for (uint32_t j = 0; j < ch0; ++j)
            cout[cex + j] = make_uint2(u0[0] + j * kChunkVsets, min(kChunkVsets, cnt[0] - j * kChunkVsets));
        for (uint32_t j = 0; j < ch1; ++j)
            cout[cex + ch0 + j] = make_uint2(u0[1] + j * kChunkVsets, min(kChunkVsets, cnt[1] - j * kChunkVsets));
        ltot += __builtin_amdgcn_readfirstlane(lsum);
        ctot += __builtin_amdgcn_readfirstlane(csum);
    }
}

// head != 0: launched right after the init kernel, for level 1 (its inputs are the root's
// rset and the reset state, so level_begin's counter reads are skipped).
__global__ void __launch_bounds__(kFusedBlock) bfs_fused_kernel(GraphView g, BfsArgs a, unsigned Lk, int head) {
    const unsigned Lin = level_of(a.ctrl, a.levelFromCtrl, Lk);
    __shared__ unsigned pre[kMaxSegs], cpre[kMaxSegs];
    __shared__ uint32_t sPf[kFuseTile], sQ0[kFuseTile], sFb[kFuseTile], sU0[kFuseTile], sUc[kFuseTile];
    __shared__ unsigned cntr, tmp32[16];
    __shared__ unsigned long long tmp64[16];
    Ctrl* c = a.ctrl;
    const uint32_t segs = a.sv.segs;
    const unsigned t = threadIdx.x, nb = blockDim.x;
    DirectionParams dp;
    dp.direction = a.direction;
    dp.canBottomUp = g.csrPtr != nullptr;
    dp.alpha = a.alpha;
    dp.beta = a.beta;
    dp.totalEdges = g.nnz;
    dp.totalItems = double(g.n);
    bool bottomUp = false;
    unsigned total = 0, ctotal = 0, Le = 0;
    LevelState cur, in;  // state of the level about to be expanded, and of the one before
    if (head) {
        // level 1: frontier {src}; level_begin's state, direction rule and list size
        const uint32_t r = a.src >> 3;
        const uint32_t u0 = g.rsetUnit[r], u1 = g.rsetUnit[r + 1];
        const unsigned long long deg = g.csrPtr ? g.csrPtr[a.src + 1] - g.csrPtr[a.src] : 0ull;
        const uint32_t ch = rset_chunks(u1 - u0);
        total = ch ? 0u : u1 - u0;
        ctotal = ch;
        Le = 1;
        cur = LevelState{0, 1, 1, 1, 0, 0, deg, 0ull};
        bottomUp = a.direction == 0 && dp.canBottomUp &&
                   double(deg) * a.alpha > double(g.nnz - min(deg, g.nnz));
        if (bottomUp || total + ctotal * kChunkVsets > fuse_vsets(a.fuseQuads)) return;
    } else {
        if (!level_begin(c, a.sc, segs, Lin, dp, bottomUp, pre, &total, cpre, &ctotal, &Le, &cur, &in)) return;
        // too big or bottom-up: the expand kernel enqueued as Lin takes it (its state slot
        // written above is rewritten there from the same inputs)
        const unsigned scale = cur.prevNf < in.prevNf ? kFuseShrinkScale : 1u;
        if (bottomUp || total + ctotal * kChunkVsets > fuse_vsets(scale * a.fuseQuads)) return;
        if (t == 0 && !c->tailFrom) c->tailFrom = Lin;
    }
    uint32_t* lists = (Lin & 1) ? a.list1 : a.list0;  // read first, rewritten when handing back
    uint2* chunks = (Lin & 1) ? a.chunk1 : a.chunk0;
    unsigned long long* F = (Lin & 1) ? a.front1 : a.front0;  // frontier bitmap, in and out
    const FusedInfo fi{sPf, sQ0, sFb, sU0, sUc};
    FusedQueue qc{pre, reinterpret_cast<uint32_t*>(a.nextB)};
    FusedQueue qn{cpre, reinterpret_cast<uint32_t*>(a.nextB) + g.R};
    unsigned nq;
    if (head) {
        if (t == 0) pre[0] = a.src >> 3;  // read below by this thread only (entry 0)
        nq = 1;
    } else {
        // list entry t and chunk entry t (at most one each): the first vset of an rset
        uint32_t cand0 = kNoRset, cand1 = kNoRset;
        if (t < total) {
            const uint32_t v = list_entry(lists, a.sv, pre, t);
            const uint32_t r = g.unitRset[v], rp = v ? g.unitRset[v - 1] : kNoRset;
            if (rp != r) cand0 = r;
        }
        if (t < ctotal) {
            const uint2 ch = chunk_entry(chunks, a.sv, cpre, t);
            const uint32_t r = g.unitRset[ch.x], rp = ch.x ? g.unitRset[ch.x - 1] : kNoRset;
            if (rp != r) cand1 = r;
        }
        if (t == 0) cntr = 0;
        __syncthreads();  // the list prefixes are read: they become the queues
        if (cand0 != kNoRset) qc.lds[atomicAdd(&cntr, 1u)] = cand0;  // <= 1024 + 64 entries
        if (cand1 != kNoRset) qc.lds[atomicAdd(&cntr, 1u)] = cand1;
        __syncthreads();
        nq = __builtin_amdgcn_readfirstlane(cntr);
    }
    uint32_t* vis32 = reinterpret_cast<uint32_t*>(a.visited);
    uint32_t* f32 = reinterpret_cast<uint32_t*>(F);
    unsigned Lcur = Le;
    LevelState prev = cur;  // state of level Lcur - 1 (from the second level on)
    unsigned long long prevEdges = 0;  // degree sum of level Lcur's frontier
    for (unsigned it = 0;; ++it) {
        const unsigned T = nq <= kFuseTile ? fused_queue_info(g, qc.lds, nq, F, fi, tmp32) : 0xFFFFFFFFu;
        if (it > 0) {
            // the entering level always runs here; a later one only while small and top-down
            const unsigned cap = cur.prevNf < prev.prevNf ? kFuseShrinkScale * a.fuseQuads : a.fuseQuads;
            bool stop = nq > kFuseTile || T > cap || a.direction == 2;
            if (!stop && a.direction == 0 && g.csrPtr) {
                const double mu = double(g.nnz - min(prev.cumEdges + prevEdges, g.nnz));
                stop = double(prevEdges) * a.alpha > mu && cur.prevNf > prev.prevNf;
            }
            if (stop) {
                // level Lcur for the expand kernel enqueued as Lin: its lists and counters
                // and the state of level Lcur - 1 with the shift (the frontier is in F)
                unsigned ltot, ctot;
                fused_write_lists(g, qc, nq, nq <= kFuseTile ? &fi : nullptr, lists, chunks, tmp32, ltot, ctot);
                for (unsigned i = kFuseTile + t; i < nq; i += nb) qc.release(i);
                if (t == 0) {
                    LevelState o = prev;
                    o.shift = Lcur - Lin;
                    c->state[(Lin - 1) & 1] = o;
                }
                const size_t base = size_t(Lin % 3) * segs;
                for (uint32_t s = t; s < segs; s += nb) {
                    a.sc.count[base + s] = s ? 0u : ltot;
                    a.sc.chunks[base + s] = s ? 0u : ctot;
                    a.sc.news[base + s] = s ? 0u : cur.prevNf;
                    a.sc.edges[base + s] = s ? 0ull : prevEdges;
                }
                return;
            }
        }
        // the queue's frontier bytes are in LDS: its words take this level's discoveries
        for (unsigned i = t; i < nq; i += nb) F[qc.lds[i] >> 3] = 0ull;
        if (t == 0) cntr = 0;
        __syncthreads();
        unsigned news = 0;
        unsigned long long edges = 0;
        for (unsigned t0 = 0; t0 < T; t0 += nb * kFuseQpl) {
            uint32_t row[4 * kFuseQpl], w[4 * kFuseQpl];
            unsigned hit = 0;  // bit s: slice s hits the frontier
#pragma unroll
            for (int h = 0; h < kFuseQpl; ++h) {
                const unsigned tt = t0 + h * nb + t;
                uint4 r4 = make_uint4(0, 0, 0, 0);
                uint32_t m = 0;
                if (tt < T) {
                    unsigned o = 0;  // the last entry whose prefix is <= tt
#pragma unroll
                    for (unsigned step = kFuseTile / 2; step > 0; step >>= 1)
                        if (o + step < nq && sPf[o + step] <= tt) o += step;
                    const uint32_t q = sQ0[o] + (tt - sPf[o]);
                    r4 = stream_load_quad(g.rows, q);
                    m = stream_load_u32(g.masks, q) & sFb[o];
                }
                row[4 * h + 0] = r4.x;
                row[4 * h + 1] = r4.y;
                row[4 * h + 2] = r4.z;
                row[4 * h + 3] = r4.w;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if ((m >> (8 * j)) & 0xFFu) hit |= 1u << (4 * h + j);
            }
            // not yet visited (a stale word only shows fewer bits: the atomic decides)
#pragma unroll
            for (int s = 0; s < 4 * kFuseQpl; ++s) w[s] = ((hit >> s) & 1u) ? vis32[row[s] >> 5] : 0u;
#pragma unroll
            for (int s = 0; s < 4 * kFuseQpl; ++s)
                if ((w[s] >> (row[s] & 31)) & 1u) hit &= ~(1u << s);
#pragma unroll
            for (int s = 0; s < 4 * kFuseQpl; ++s)
                w[s] = ((hit >> s) & 1u) ? atomicOr(&vis32[row[s] >> 5], 1u << (row[s] & 31)) : 0u;
#pragma unroll
            for (int s = 0; s < 4 * kFuseQpl; ++s)
                if ((w[s] >> (row[s] & 31)) & 1u) hit &= ~(1u << s);
            // the rows claimed here
#pragma unroll
            for (int s = 0; s < 4 * kFuseQpl; ++s) {
                if ((hit >> s) & 1u) {
                    a.levels[row[s]] = int32_t(Lcur);
                    w[s] = atomicOr(&f32[row[s] >> 5], 1u << (row[s] & 31));
                    if (g.csrPtr) {  // both list bounds in one 8-byte load
                        const u32x2_a4 d = *reinterpret_cast<const u32x2_a4*>(g.csrPtr + row[s]);
                        edges += d.y - d.x;
                    }
                    news += 1;
                }
            }
#pragma unroll
            for (int s = 0; s < 4 * kFuseQpl; ++s)
                if (((hit >> s) & 1u) && ((w[s] >> (row[s] & 24)) & 0xFFu) == 0u)
                    qn.put(atomicAdd(&cntr, 1u), row[s] >> 3);  // first vertex of its rset
        }
        const unsigned nf = __builtin_amdgcn_readfirstlane(unsigned(block_sum_u64(news, tmp64)));
        const unsigned long long eSum = block_sum_u64(edges, tmp64);  // the sums' barriers follow every append
        const unsigned long long mf = (unsigned long long)__builtin_amdgcn_readfirstlane(unsigned(eSum >> 32)) << 32 |
                                      __builtin_amdgcn_readfirstlane(unsigned(eSum));
        const unsigned nNew = __builtin_amdgcn_readfirstlane(cntr);
        if (nf == 0) {
            if (t == 0) {
                c->levels = Lcur;
                c->fusedLevels = Lcur - (Lin - 1);
                c->tdLevels = cur.td;
                c->buLevels = cur.bu;
                c->visited = cur.visited;
                c->visitedEdges = cur.cumEdges;
                c->buMask = cur.buMask;
                __hip_atomic_store(&c->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
        prev = cur;
        cur.dir = 0;
        cur.prevNf = nf;
        cur.visited = prev.visited + nf;
        cur.td = prev.td + 1;
        cur.cumEdges = prev.cumEdges + mf;
        prevEdges = mf;
        ++Lcur;
        const FusedQueue tq = qc;
        qc = qn;
        qn = tq;
        nq = nNew;
    }
}

// Two launch-time knobs, both chosen by the autotuner: QPL = quads each lane streams per
// top-down step (the slice tile a wave holds in flight: 64 x QPL quads = 256 x QPL
// slices), BUG = bitmap words a bottom-up wave stages per group (vertices whose first
// in-neighbour is probed together, BUG per lane).
template <int QPL, int BUG>
__global__ void __launch_bounds__(kProducerBlock) BLEST_RESIDENT_GRID bfs_expand_kernel(GraphView g, BfsArgs a, unsigned Lk) {
    const unsigned L = level_of(a.ctrl, a.levelFromCtrl, Lk);
    __shared__ unsigned pre[kMaxSegs];
    __shared__ unsigned long long tmp64[16];
    DirectionParams dp;
    dp.direction = a.direction;
    dp.canBottomUp = g.csrPtr != nullptr;
    dp.alpha = a.alpha;
    dp.beta = a.beta;
    dp.totalEdges = g.nnz;
    dp.totalItems = double(g.n);
    Ctrl* c = a.ctrl;
    const uint32_t segs = a.sv.segs;
    bool bottomUp = false;
    __shared__ unsigned cpre[kMaxSegs];
    unsigned total = 0, ctotal = 0;  // list entries / chunks of the level (top-down)
    unsigned Le;                     // the level's distance
    LevelState st, in;
    if (!level_begin(c, a.sc, segs, L, dp, bottomUp, pre, &total, cpre, &ctotal, &Le, &st, &in)) return;
    if (a.fuseQuads && blockIdx.x == 0 && threadIdx.x == 0) {
        // where the next traversal's fused kernel should take over (the level driver's hint)
        const bool small = !bottomUp && st.prevNf < in.prevNf &&
                           total + ctotal * kChunkVsets <= fuse_vsets(kFuseShrinkScale * a.fuseQuads);
        c->tailFrom = small ? (c->tailFrom ? c->tailFrom : L) : 0u;
    }
    const unsigned lane = lane_id();
    const unsigned long long* fin = (L & 1) ? a.front1 : a.front0;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    const unsigned sOut = (L + 1) % 3;
    if (!bottomUp) {
        const uint32_t* lin = (L & 1) ? a.list1 : a.list0;
        const uint2* cin = (L & 1) ? a.chunk1 : a.chunk0;
        const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
        const unsigned nwaves = (gridDim.x * blockDim.x) / kWave;
        // The wave's lanes hold vsets (v valid where ok): their quads are streamed, the
        // union of the lanes' quad runs 64 (x2) at a time.
        auto stream = [&](uint32_t v, bool ok) {
            uint32_t q0 = 0, qc = 0, fb = 0;
            if (ok) {
                q0 = g.unitQuad[v];
                qc = g.unitQuad[v + 1] - q0;
                fb = front_byte(fin, g.unitRset[v]);
                if (!fb) qc = 0;
            }
            const unsigned incl = wave_inclusive_scan(qc);
            const unsigned excl = incl - qc;
            const unsigned tq = __shfl(incl, kWave - 1, kWave);
            // QPL quads per lane per step, so QPL row/mask loads are in flight
            for (unsigned t0 = 0; t0 < tq; t0 += QPL * kWave) {
                uint4 r4[QPL];
                uint32_t m[QPL];
#pragma unroll
                for (int h = 0; h < QPL; ++h) {
                    const unsigned t = t0 + h * kWave + lane;
                    const unsigned owner = wave_owner(excl, min(t, tq - 1));
                    const uint32_t oq0 = __shfl(q0, owner, kWave);
                    const unsigned oex = __shfl(excl, owner, kWave);
                    const uint32_t of = __shfl(fb, owner, kWave);
                    r4[h] = make_uint4(0, 0, 0, 0);
                    m[h] = 0;
                    if (t < tq) {
                        const uint32_t q = oq0 + (t - oex);
                        r4[h] = stream_load_quad(g.rows, q);
                        m[h] = stream_load_u32(g.masks, q) & (of * 0x01010101u);
                    }
                }
#pragma unroll
                for (int h = 0; h < QPL; ++h) {
                    expand_slice(r4[h].x, m[h] & 0xFF, a.visited, a.nextB);
                    expand_slice(r4[h].y, (m[h] >> 8) & 0xFF, a.visited, a.nextB);
                    expand_slice(r4[h].z, (m[h] >> 16) & 0xFF, a.visited, a.nextB);
                    expand_slice(r4[h].w, (m[h] >> 24) & 0xFF, a.visited, a.nextB);
                }
            }
        };
        // m list entries (or chunk slots) per wave step, from the level's whole work
        // (list entries + kChunkVsets slots per chunk): a short list of large vsets (the second
        // level of an R-MAT traversal: a few thousand vsets of ~100 slices) is spread over
        // as many waves as it has entries instead of 64 entries per wave, and a long one
        // takes as few steps per wave as 64 vsets per step allow
        const unsigned cslots = ctotal * kChunkVsets;
        const unsigned target = (total + cslots + nwaves - 1) / nwaves;
        const unsigned m = min(unsigned(kWave), max(1u, target));
        for (unsigned base = wave * m; base < total; base += nwaves * m) {
            const unsigned e = base + lane;
            const bool ok = lane < m && e < total;
            stream(ok ? list_entry(lin, a.sv, pre, e) : 0u, ok);
        }
        // then the chunks (<= kChunkVsets consecutive vsets of one rset), seen as
        // kChunkVsets slots each and
        // enumerated offset-major (slot 0 of every chunk, then slot 1, ...), m slots per
        // wave step like the list: a step's lanes hold vsets of different chunks, so their
        // rows (and `next` atomics) are spread as a list step's are, instead of 64
        // consecutive vsets of one hub whose rows crowd into the same bitmap words; the
        // round robin over the waves continues after the list's steps
        const unsigned lsteps = (total + m - 1) / m;
        for (unsigned base = ((wave + nwaves - lsteps % nwaves) % nwaves) * m; base < cslots; base += nwaves * m) {
            const unsigned e = base + lane;
            bool ok = lane < m && e < cslots;
            uint32_t v = 0;
            if (ok) {
                const uint2 ch = chunk_entry(cin, a.sv, cpre, e % ctotal);
                const uint32_t off = e / ctotal;
                ok = off < ch.y;
                v = ch.x + off;
            }
            stream(v, ok);
        }
        return;
    }
    // bottom-up over this workgroup's segment.  Wave wib takes the words ws + wib,
    // ws + wib + wpb, ... in groups of kBuGroup; lane l owns vertex 64w + l of every
    // word of the group and scans the in-neighbours of its open vertices one after
    // another, 1, 2, then kBuBatch entries per step (a step's loads issued together;
    // most vertices of a bottom-up level hit at their first entry), independently of
    // the other lanes: a wave runs as long as its busiest lane, not as the sum over its
    // words of each word's slowest lane.  A scan still without a hit after kBuBudget
    // entries is handed, from where it stopped, to a pass in which a whole wave scans
    // the rest 64 entries at a time, so a long unsuccessful scan does not hold one lane
    // for the whole level.  (A fully wave-cooperative scan, lanes dealt out over the
    // word's open vertices every step, measured slower: its per-step shuffles cost as
    // much as a memory round trip.)
    constexpr int kBuGroup = BUG, kBuBatch = 4;
    constexpr uint32_t kBuBudget = 32;
    constexpr unsigned kHeavyCap = 1024;
    __shared__ uint32_t heavyU[kHeavyCap], heavyP[kHeavyCap];
    __shared__ unsigned heavyN;
    if (threadIdx.x == 0) heavyN = 0;
    __syncthreads();
    const uint32_t ws = blockIdx.x * a.sv.wps, we = min(g.W, ws + a.sv.wps);
    // frontier tests read the 32-bit half of the bitmap word that holds the bit
    const uint32_t* fin32 = reinterpret_cast<const uint32_t*>(fin);
    // wave-uniform (readfirstlane): the group's word indices live in scalar registers
    const unsigned wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), wpb = blockDim.x / kWave;
    unsigned long long news = 0, edges = 0;
    for (uint32_t w0 = ws + wib; w0 < we; w0 += wpb * kBuGroup) {
        uint32_t pb[kBuGroup], pe[kBuGroup];
        unsigned cand = 0;   // bit k: the vertex of word k is open and has in-neighbours
        unsigned openm = 0;  // bit k: the vertex of word k is open
#pragma unroll
        for (int k = 0; k < kBuGroup; ++k) {
            const uint32_t w = w0 + k * wpb;
            // a vertex without in-neighbours is never found bottom-up: closed like a visited one
            const unsigned long long vis = w < we ? (a.visited[w] | (g.noIn ? g.noIn[w] : 0ull)) : ~0ull;
            const uint32_t u = w * 64 + lane;
            const bool open = w < we && u < g.n && !((vis >> lane) & 1ull);
            if (open) openm |= 1u << k;
            pb[k] = open ? g.csrPtr[u] : 0u;
            pe[k] = open ? g.csrPtr[u + 1] : 0u;
            if (pe[k] > pb[k]) cand |= 1u << k;
        }
        unsigned found = 0;
        // The first in-neighbour of every open vertex of the group at once (all index
        // loads, then all frontier words): most vertices of a bottom-up level hit there,
        // so the group costs one chain of dependent loads, not one per word.  With the
        // lists sorted by degree the first entries are also in g.probe, read coalesced
        // (lane l of word w loads entry 64w + l).  A vertex still open afterwards goes on
        // from its second entry (pb stays the list start: its degree is counted when it
        // is found).
        {
            // the probe array comes with the no-in-neighbour bitmap, so there an open
            // vertex has an in-neighbour: its probe goes out with its list bounds, one
            // round trip before they return
            const unsigned first = g.probe ? openm : cand;
            uint32_t nb0[kBuGroup], fw0[kBuGroup];
#pragma unroll
            for (int j = 0; j < kBuGroup; ++j)
                nb0[j] = ((first >> j) & 1u) ? (g.probe ? g.probe[(w0 + j * wpb) * 64 + lane] : g.csrIdx[pb[j]]) : 0u;
#pragma unroll
            for (int j = 0; j < kBuGroup; ++j) fw0[j] = ((first >> j) & 1u) ? fin32[nb0[j] >> 5] : 0u;
#pragma unroll
            for (int j = 0; j < kBuGroup; ++j) {
                if ((cand >> j) & 1u) {
                    const bool hit0 = (fw0[j] >> (nb0[j] & 31)) & 1u;
                    if (hit0) found |= 1u << j;
                    if (hit0 || pb[j] + 1 >= pe[j]) cand &= ~(1u << j);
                }
            }
        }
        int k = cand ? __builtin_ctz(cand) : kBuGroup;
        uint32_t p = pick(pb, k) + 1, p1 = pick(pe, k);
        uint32_t budget = p + kBuBudget;
        unsigned step = 2;
        while (k < kBuGroup) {  // p < p1 here
            const uint32_t lim = min(p + step, p1);
            uint32_t nb[kBuBatch], fw[kBuBatch];
#pragma unroll
            for (int i = 0; i < kBuBatch; ++i) nb[i] = p + i < lim ? g.csrIdx[p + i] : 0u;
#pragma unroll
            for (int i = 0; i < kBuBatch; ++i) fw[i] = p + i < lim ? fin32[nb[i] >> 5] : 0u;
            bool hit = false;
#pragma unroll
            for (int i = 0; i < kBuBatch; ++i) hit |= (fw[i] >> (nb[i] & 31)) & 1u;
            p = lim;
            step = min(2 * step, unsigned(kBuBatch));
            bool next = hit || p >= p1;
            if (!next && p >= budget) {
                const unsigned slot = atomicAdd(&heavyN, 1u);
                if (slot < kHeavyCap) {
                    heavyU[slot] = (w0 + k * wpb) * 64 + lane;
                    heavyP[slot] = p;
                    next = true;
                } else {
                    budget = p1;  // queue full: this lane finishes the scan itself
                }
            }
            if (next) {
                if (hit) found |= 1u << k;
                cand &= ~(1u << k);
                k = cand ? __builtin_ctz(cand) : kBuGroup;
                p = pick(pb, k) + 1;
                p1 = pick(pe, k);
                budget = p + kBuBudget;
                step = 2;
            }
        }
#pragma unroll
        for (int k = 0; k < kBuGroup; ++k) {
            const uint32_t w = w0 + k * wpb;
            const bool f = (found >> k) & 1u;
            const unsigned long long nw = __ballot(f);
            if (w < we) {
                if (lane == 0) {
                    if (nw) atomicOr(&a.visited[w], nw);  // result unused: no round trip
                    fout[w] = nw;
                }
                if (f) {
                    a.levels[w * 64 + lane] = int32_t(Le);
                    news += 1;
                    edges += pe[k] - pb[k];
                }
            }
        }
    }
    __syncthreads();  // the queue is complete and every word's first store precedes the ORs below
    const unsigned nh = min(heavyN, kHeavyCap);
    for (unsigned h = wib; h < nh; h += wpb) {
        const uint32_t u = heavyU[h];
        const uint32_t p0 = g.csrPtr[u], p1 = g.csrPtr[u + 1];
        bool hit = false;
        for (uint32_t q0 = heavyP[h]; q0 < p1 && !hit; q0 += kWave) {  // wave-uniform
            const uint32_t q = q0 + lane;
            hit = __ballot(q < p1 && test_bit(fin, g.csrIdx[min(q, p1 - 1)])) != 0ull;
        }
        if (hit && lane == 0) {
            const unsigned long long bit = 1ull << (u & 63);
            atomicOr(&a.visited[u >> 6], bit);
            atomicOr(&fout[u >> 6], bit);
            a.levels[u] = int32_t(Le);
            news += 1;
            edges += p1 - p0;
        }
    }
    news = block_sum_u64(news, tmp64);
    edges = block_sum_u64(edges, tmp64);
    if (threadIdx.x == 0) {
        a.sc.news[size_t(sOut) * segs + blockIdx.x] = unsigned(news);
        a.sc.edges[size_t(sOut) * segs + blockIdx.x] = edges;
    }
}

// One pass, thread per rset: the block's segment is walked in chunks of blockDim/8
// words (one rset per thread, the 8 rsets of a word in one wave).  Top-down levels
// fold `next` into visited / frontier (the word's first thread), write the new
// vertices' levels and sum their degrees; both directions then emit the active rsets'
// vsets (one block scan per chunk for both lists; each rset writes its own entries).
// A chunk's loads are as independent as the data allows: its discovery bytes were
// loaded during the previous chunk, then come the active rsets' vset ranges, then the
// degrees of all of a thread's eight new vertices at once (every load in flight before
// the first is used; one round trip per vertex made the dense levels of deep graphs cost
// ~10 us per chunk).
constexpr int kDegBatch = 8;  // degree loads in flight per thread: all of its vertices in a chunk
__global__ void __launch_bounds__(kProducerBlock) BLEST_RESIDENT_GRID bfs_update_kernel(GraphView g, BfsArgs a, unsigned Lk) {
    const unsigned L = level_of(a.ctrl, a.levelFromCtrl, Lk);
    __shared__ unsigned tmp[32];
    __shared__ unsigned long long tmp64[16];
    // the chunk's new-vertex words, by chunk parity: a chunk's levels are written after its
    // lists, while faster waves may already store the next chunk's words
    __shared__ unsigned long long sWord[2][kProducerBlock / 8];
    Ctrl* c = a.ctrl;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    const uint32_t segs = a.sv.segs;
    const unsigned sOut = (L + 1) % 3;
    const uint32_t ws = blockIdx.x * a.sv.wps, we = min(g.W, ws + a.sv.wps);
    const unsigned t = threadIdx.x, b = t & 7, nb = blockDim.x;
    // one batch of independent loads: the flags, the segment's list bases and both
    // candidate words of the first chunk (next for top-down, fout for bottom-up)
    const unsigned done = ld_relaxed(&c->done);
    const unsigned dir = c->state[L & 1].dir, shift = c->state[L & 1].shift, prevNf = c->state[L & 1].prevNf;
    const uint32_t listBase = a.sv.segBase[blockIdx.x], chunkBase = a.sv.chunkBase[blockIdx.x];
    const uint32_t w1 = min(ws + t / 8, g.W - 1);
    unsigned long long* nb64 = reinterpret_cast<unsigned long long*>(a.nextB);
    const unsigned long long firstNB = nb64[size_t(w1) * 8 + b], firstFout = fout[w1];
    // a bottom-up level's discovery counts are complete (its expand kernel wrote them)
    unsigned long long nfPart = 0;
    if (dir != 0)
        for (uint32_t s = t; s < segs; s += nb) nfPart += a.sc.news[size_t(sOut) * segs + s];
    if (done) return;
    const bool bottomUp = dir != 0;
    const unsigned Le = L + shift;  // the level's distance
    uint32_t* out = ((L & 1) ? a.list0 : a.list1) + listBase;
    uint2* cout = ((L & 1) ? a.chunk0 : a.chunk1) + chunkBase;
    const bool canBU = g.csrPtr != nullptr;
    if (bottomUp) {
        // The lists serve a top-down next level only.  Its direction is level_begin's rule
        // on the same inputs (this level's state and discovery count): when it stays
        // bottom-up, or nothing was discovered, no list is written.
        const unsigned long long nf = block_sum_u64(nfPart, tmp64);
        const bool nextBottomUp = (a.direction == 0 && canBU)
                                      ? !(double(nf) * a.beta < double(g.n) && nf < prevNf)
                                      : (a.direction == 2 && canBU);
        if (nf == 0 || nextBottomUp) {
            if (t == 0) {
                a.sc.count[size_t(sOut) * segs + blockIdx.x] = 0u;
                a.sc.chunks[size_t(sOut) * segs + blockIdx.x] = 0u;
            }
            return;
        }
    }
    unsigned long long news = 0, edges = 0;
    unsigned written = 0, cwritten = 0, pass = 0;
    // the thread's candidate (next bytes, or the frontier word) of the chunk about to run:
    // each chunk loads the next chunk's before its own work (one load in flight throughout)
    unsigned long long pre = ws + t / 8 < we ? (bottomUp ? firstFout : firstNB) : 0ull;
    for (uint32_t c0 = ws; c0 < we; c0 += nb / 8, ++pass) {
        unsigned long long* sw = sWord[pass & 1];
        const uint32_t w = c0 + t / 8, r = w * 8 + b, wn = w + nb / 8;
        const unsigned long long cur = pre;
        pre = wn < we ? (bottomUp ? fout[wn] : nb64[size_t(wn) * 8 + b]) : 0ull;
        uint32_t start = 0, cnt = 0, chunks = 0;
        bool any = false;
        unsigned long long word = 0;
        if (!bottomUp) {
            // the word's discoveries: thread b holds bytes [64w + 8b, 64w + 8b + 8) of
            // nextB, turns them into bits 8b..8b+7, and the word's 8 threads (aligned lanes
            // of one wave) OR them together; each clears the bytes it read
            const unsigned long long raw = cur;
            unsigned m8 = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) m8 |= ((raw >> (8 * i)) & 0xFFull) ? (1u << i) : 0u;
            word = (unsigned long long)m8 << (8 * b);
            word |= __shfl_xor(word, 1, kWave);
            word |= __shfl_xor(word, 2, kWave);
            word |= __shfl_xor(word, 4, kWave);
            if (raw) nb64[size_t(w) * 8 + b] = 0ull;
        }
        if (w < we) {
            if (bottomUp) {
                word = cur;
            } else if (b == 0) {
                if (word) atomicOr(&a.visited[w], word);  // result unused: no round trip
                fout[w] = word;
                sw[t / 8] = word;
            }
            const uint32_t byte = uint32_t(word >> (8 * b)) & 0xFFu;
            any = byte != 0;
            if (byte) {
                start = g.rsetUnit[r];
                cnt = g.rsetUnit[r + 1] - start;
                chunks = rset_chunks(cnt);  // a hub's rset: chunk entries, not vsets
            }
        } else if (!bottomUp && b == 0) {
            sw[t / 8] = 0ull;
        }
        // a chunk without any frontier vertex (most of them on a narrow level) needs no
        // scan, levels or list: one block-wide vote instead
        if (!__syncthreads_or(any)) continue;
        unsigned total, cex, ctotal;
        const unsigned ex = block_exclusive_scan2(chunks ? 0u : cnt, chunks, tmp, total, cex, ctotal);  // publishes sw
        // an rset listed per vset (at most kChunkVsets) writes its own entries: no owner
        // search (a thread's entries lay a workgroup apart, so each was a search of
        // dependent LDS reads) and no barrier
        if (!chunks)
            for (uint32_t j = 0; j < cnt; ++j) out[written + ex + j] = start + j;
        written += total;
        for (uint32_t j = 0; j < chunks; ++j)
            cout[cwritten + cex + j] = make_uint2(start + j * kChunkVsets, min(kChunkVsets, cnt - j * kChunkVsets));
        cwritten += ctotal;
        if (!bottomUp) {
            // the new vertices' levels and degrees, lane per vertex (coalesced): the
            // chunk's 64 words hold 8 x blockDim vertices, word wv + i x (waves per block)
            // for the wave's i-th step, taken kDegBatch steps at a time
            const unsigned wv = __builtin_amdgcn_readfirstlane(t / kWave), lane = t % kWave;
#pragma unroll
            for (int i0 = 0; i0 < 8; i0 += kDegBatch) {
                uint32_t lo[kDegBatch], hi[kDegBatch];
#pragma unroll
                for (int j = 0; j < kDegBatch; ++j) {
                    const unsigned wi = wv + (i0 + j) * (nb / kWave);  // word within the chunk
                    const uint32_t v = (c0 + wi) * 64 + lane;
                    lo[j] = hi[j] = 0u;
                    if ((sw[wi] >> lane) & 1ull) {
                        a.levels[v] = int32_t(Le);
                        news += 1;
                        if (canBU) {
                            lo[j] = g.csrPtr[v];
                            hi[j] = g.csrPtr[v + 1];
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kDegBatch; ++j) edges += hi[j] - lo[j];
            }
        }
    }
    if (!bottomUp) {
        news = block_sum_u64(news, tmp64);
        edges = block_sum_u64(edges, tmp64);
    }
    if (t == 0) {
        a.sc.count[size_t(sOut) * segs + blockIdx.x] = written;
        a.sc.chunks[size_t(sOut) * segs + blockIdx.x] = cwritten;
        if (!bottomUp) {
            a.sc.news[size_t(sOut) * segs + blockIdx.x] = unsigned(news);
            a.sc.edges[size_t(sOut) * segs + blockIdx.x] = edges;
        }
    }
}

// Unreached vertices get -1 (lane per vertex: coalesced stores).
__global__ void __launch_bounds__(kProducerBlock) bfs_final_kernel(GraphView g, BfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    const unsigned lane = lane_id();
    for (uint32_t w = wave; w < g.W; w += nwaves) {
        const unsigned long long vis = a.visited[w];
        const uint32_t u = w * 64 + lane;
        if (u < g.n && !((vis >> lane) & 1ull)) a.levels[u] = -1;
    }
}

__global__ void unpermute_levels_kernel(int32_t* out, const int32_t* in, const uint32_t* oldToNew, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[oldToNew[i]];
}

}  // namespace dev

// The expand kernel instantiation for the graph's tile knobs (validated at creation).
using ExpandKernel = void (*)(GraphView, dev::BfsArgs, unsigned);
static ExpandKernel expand_kernel_for(const DeviceGraph& g) {
    const bool g4 = g.opts.bu_group == 4;
    switch (g.opts.td_quads) {
        case 1: return g4 ? dev::bfs_expand_kernel<1, 4> : dev::bfs_expand_kernel<1, 8>;
        case 4: return g4 ? dev::bfs_expand_kernel<4, 4> : dev::bfs_expand_kernel<4, 8>;
        default: return g4 ? dev::bfs_expand_kernel<2, 4> : dev::bfs_expand_kernel<2, 8>;
    }
}

void bfs_launch(Query& q, uint32_t src, bool trace) {
    DeviceGraph& g = *q.g;
    dev::BfsArgs a;
    a.levels = q.levels;
    a.visited = q.visited;
    a.front0 = q.front0;
    a.front1 = q.front1;
    a.nextB = q.nextB;
    a.list0 = q.list0;
    a.list1 = q.list1;
    a.chunk0 = q.chunk0;
    a.chunk1 = q.chunk1;
    a.ctrl = q.ctrl;
    a.sv = g.seg_view();
    a.sc = q.seg_counters();
    a.src = src;
    a.direction = g.hasCsr ? g.opts.direction : 1;
    a.alpha = g.opts.alpha;
    a.beta = g.opts.beta;
    a.levelFromCtrl = 0;
    // a trace times every level's own kernels
    a.fuseQuads = (trace || g.opts.fuse_quads < 0) ? 0u : unsigned(g.opts.fuse_quads);
    dev::BfsArgs ar = a;
    ar.levelFromCtrl = 1;
    const GraphView gv = g.view();
    const dim3 grid = level_grid(g), block(dev::kProducerBlock);
    const ExpandKernel expand = expand_kernel_for(g);
    BLEST_HIP(hipEventRecord(q.ev0, q.stream));
    hipLaunchKernelGGL(dev::bfs_init_kernel, grid, block, 0, q.stream, gv, a);
    // small levels between chunks of level kernels: from the root, and after every chunk
    const bool fused = a.fuseQuads != 0;
    if (fused) hipLaunchKernelGGL(dev::bfs_fused_kernel, dim3(1), dim3(dev::kFusedBlock), 0, q.stream, gv, a, 1u, 1);
    if (trace) {
        size_t next = 0;
        auto mark = [&] {
            if (next == q.traceEvents.size()) {
                hipEvent_t e;
                BLEST_HIP(hipEventCreate(&e));
                q.traceEvents.push_back(e);
            }
            BLEST_HIP(hipEventRecord(q.traceEvents[next++], q.stream));
        };
        run_levels(
            q,
            [&](unsigned L) {
                mark();
                hipLaunchKernelGGL(expand, grid, block, 0, q.stream, gv, a, L);
                mark();
                hipLaunchKernelGGL(dev::bfs_update_kernel, grid, block, 0, q.stream, gv, a, L);
                mark();
            },
            [&] { hipLaunchKernelGGL(dev::bfs_final_kernel, grid, block, 0, q.stream, gv, a); });
        BLEST_HIP(hipGetLastError());
        BLEST_HIP(hipEventRecord(q.ev1, q.stream));
        q.lastSource = src;
        q.pending = true;
        return;
    }
    run_levels(
        q,
        [&](unsigned L) {
            hipLaunchKernelGGL(expand, grid, block, 0, q.stream, gv, a, L);
            hipLaunchKernelGGL(dev::bfs_update_kernel, grid, block, 0, q.stream, gv, a, L);
        },
        [&] { hipLaunchKernelGGL(dev::bfs_final_kernel, grid, block, 0, q.stream, gv, a); },
        &q.graphs[kGraphBfs],
        [&](unsigned k) {
            hipLaunchKernelGGL(expand, grid, block, 0, q.stream, gv, ar, k);
            hipLaunchKernelGGL(dev::bfs_update_kernel, grid, block, 0, q.stream, gv, ar, k);
        },
        launch_key(a.alpha, a.beta, a.direction,
                   g.opts.td_quads | (g.opts.bu_group << 8) | (uint64_t(a.fuseQuads) << 16)),
        fused ? std::function<void(unsigned)>([&](unsigned L) {
            hipLaunchKernelGGL(dev::bfs_fused_kernel, dim3(1), dim3(dev::kFusedBlock), 0, q.stream, gv, a, L, 0);
        })
              : nullptr,
        fused ? std::function<void(unsigned)>([&](unsigned k) {
            hipLaunchKernelGGL(dev::bfs_fused_kernel, dim3(1), dim3(dev::kFusedBlock), 0, q.stream, gv, ar, k, 0);
        })
              : nullptr);
    BLEST_HIP(hipGetLastError());
    BLEST_HIP(hipEventRecord(q.ev1, q.stream));
    q.lastSource = src;
    q.pending = true;
}

void bfs_finish(Query& q, int32_t* hostLevels, blest_bfs_stats* st) {
    DeviceGraph& g = *q.g;
    BLEST_HIP(hipStreamSynchronize(q.stream));
    q.pending = false;
    dev::Ctrl c;
    BLEST_HIP(hipMemcpyAsync(&c, q.ctrl, sizeof(c), hipMemcpyDeviceToHost, q.stream));
    BLEST_HIP(hipStreamSynchronize(q.stream));
    // level kernels the next traversal enqueues before the fused kernel takes over: up to
    // the final run of small shrinking levels, else all this traversal used
    q.levelHint = c.tailFrom ? c.tailFrom - 1 : c.levels - c.fusedLevels;
    float ms = 0;
    BLEST_HIP(hipEventElapsedTime(&ms, q.ev0, q.ev1));
    if (st) {
        st->kernel_ms = ms;
        st->total_levels = c.levels;
        st->no_visited = c.visited;
        st->td_levels = c.tdLevels;
        st->bu_levels = c.buLevels;
        st->edges_traversed_component = c.visitedEdges / 2;
        st->bottom_up_mask = c.buMask;
    }
    if (hostLevels) {
        const int32_t* src = q.levels;
        if (g.hasPerm) {
            hipLaunchKernelGGL(dev::unpermute_levels_kernel, dim3(1024), dim3(256), 0, q.stream, q.levelsOut,
                               q.levels, g.permOldToNew, g.n);
            BLEST_HIP(hipGetLastError());
            src = q.levelsOut;
        }
        BLEST_HIP(hipMemcpyAsync(hostLevels, src, sizeof(int32_t) * g.n, hipMemcpyDefault, q.stream));
        BLEST_HIP(hipStreamSynchronize(q.stream));
    }
}

}  // namespace blest

using namespace blest;

extern "C" {

blest_status blest_bfs(blest_query* qh, uint32_t src, int32_t* levels, blest_bfs_stats* st) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        if (levels) check_result_range(levels, sizeof(int32_t) * size_t(q.g->n), q.g->device, "levels");
        bfs_launch(q, map_source(*q.g, src));
        bfs_finish(q, levels, st);
    });
}

// The traversal runs on the query's own stream.  A caller stream, when given, is joined
// both ways: the traversal starts after the work already queued on it, and work queued
// on it afterwards (e.g. a read of blest_query_levels_device) starts after the traversal.
blest_status blest_bfs_async(blest_query* qh, uint32_t src, void* stream) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        const uint32_t s = map_source(*q.g, src);
        hipStream_t caller = static_cast<hipStream_t>(stream);
        if (caller) {
            BLEST_HIP(hipEventRecord(q.evIn, caller));
            BLEST_HIP(hipStreamWaitEvent(q.stream, q.evIn, 0));
        }
        bfs_launch(q, s);
        if (caller) BLEST_HIP(hipStreamWaitEvent(caller, q.ev1, 0));
    });
}

blest_status blest_query_sync(blest_query* qh, blest_bfs_stats* st) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        bfs_finish(q, nullptr, st);
    });
}

void* blest_query_levels_device(blest_query* qh) { return qh ? qh->q.levels : nullptr; }

blest_status blest_bfs_trace(blest_query* qh, uint32_t src, int32_t* levels, blest_bfs_stats* st,
                             blest_level_trace* trace, uint32_t cap, uint32_t* count) {
    return guarded([&] {
        if (!qh || (cap && !trace)) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        if (levels) check_result_range(levels, sizeof(int32_t) * size_t(q.g->n), q.g->device, "levels");
        bfs_launch(q, map_source(*q.g, src), true);
        blest_bfs_stats s{};
        bfs_finish(q, levels, &s);
        if (st) *st = s;
        // events per level L = 1, 2, ...: before expand(L), after it, after update(L)
        const uint32_t nl = std::min<uint32_t>(s.total_levels, uint32_t(q.traceEvents.size() / 3));
        const uint32_t m = std::min(cap, nl);
        for (uint32_t i = 0; i < m; ++i) {
            const unsigned L = i + 1;
            float e = 0, u = 0;
            BLEST_HIP(hipEventElapsedTime(&e, q.traceEvents[3 * i], q.traceEvents[3 * i + 1]));
            BLEST_HIP(hipEventElapsedTime(&u, q.traceEvents[3 * i + 1], q.traceEvents[3 * i + 2]));
            trace[i].level = L;
            trace[i].bottom_up = L <= 64 ? uint32_t((s.bottom_up_mask >> (L - 1)) & 1ull) : 0u;
            trace[i].expand_ms = e;
            trace[i].update_ms = u;
        }
        if (count) *count = m;
    });
}

}  // extern "C"
