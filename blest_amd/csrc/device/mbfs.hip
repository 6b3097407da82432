// Multi-source BFS: up to 128 independent traversals advanced together (gfx950,
// wave64).  The reference declares this plugin (MBFS/MBFSKernel.cuh:21-53) but its
// kernel file is empty; the contract is "one level array per root".
//
// Every vertex carries NW 64-bit words (NW = 1: 64 roots, NW = 2: 128 roots): bit j of
// visited[v] / frontier[v] says whether root j has reached / just reached v.  Level L is
// two stream-ordered kernels, like the single-source traversal, with the same
// segmented (atomic-free) frontier lists and counters:
//   expand  top-down: a slice (row i, mask m) of rset r contributes OR_{b in m} F[8r+b]
//             to row i: the boolean product of the slice with the 8 x 64*NW frontier
//             block of its rset.  Roots not yet at i are OR-ed into next[i].
//           bottom-up: workgroup per segment, lane per vertex: the roots that have not
//             reached v pull the OR of their in-neighbours' frontier words, stopping
//             once all are found.
//   update  top-down: lane per vertex: new = next[v]; visited |= new; levels of the new
//             (root, vertex) pairs; both directions list the vsets of rsets holding a
//             frontier vertex of any root.
// Levels are written root-major: levels[j * n + v].
#include "dgraph.hpp"
#include "level_common.hpp"

namespace blest {
namespace dev {

struct MbfsArgs {
    int32_t* levels;
    unsigned long long* vn;  // per vertex: NW visited words, then NW `next` words (one line)
    unsigned long long* front0;  // frontier words of levels past the history, by parity
    unsigned long long* front1;
    unsigned long long* hist;    // frontier words of levels 0..histLevels (n * NW each)
    uint32_t histLevels;
    unsigned long long* full;    // bit v: every root has reached v (top-down skips its record)
    unsigned long long* any;     // 2 x W: bit v of half L & 1: v is in the frontier of level L for some root
    int pull;                    // symmetric graph: bottom-up levels pull over the BVSS slices
    uint8_t* dirty;              // W bytes: word w holds a vertex whose `next` word was set this level
    uint32_t* list0;
    uint32_t* list1;
    const uint32_t* srcs;
    uint32_t k;
    Ctrl* ctrl;
    SegView sv;
    SegCounters sc;
    int direction;
    double alpha, beta;
    int levelFromCtrl;  // 1: the kernel's L argument is relative to ctrl->levelBase
};

template <int NW>
struct Bits {
    unsigned long long w[NW];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = 0;
    }
    __device__ __forceinline__ bool any() const {
        unsigned long long x = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) x |= w[i];
        return x != 0;
    }
    __device__ __forceinline__ unsigned popc() const {
        unsigned c = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) c += __popcll(w[i]);
        return c;
    }
    __device__ __forceinline__ void load(const unsigned long long* p, uint32_t v) {
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = p[size_t(v) * NW + i];
    }
    __device__ __forceinline__ void store(unsigned long long* p, uint32_t v) const {
#pragma unroll
        for (int i = 0; i < NW; ++i) p[size_t(v) * NW + i] = w[i];
    }
    __device__ __forceinline__ void loadp(const unsigned long long* p) {
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = p[i];
    }
    __device__ __forceinline__ void storep(unsigned long long* p) const {
#pragma unroll
        for (int i = 0; i < NW; ++i) p[i] = w[i];
    }
};

// A vertex's visited and `next` words sit side by side (2 x NW words, 32 B at 128 roots),
// so the random row access of a top-down slice touches one cache line, not two.
template <int NW>
__device__ __forceinline__ unsigned long long* vis_at(const MbfsArgs& a, uint32_t v) {
    return a.vn + size_t(v) * (2 * NW);
}
template <int NW>
__device__ __forceinline__ unsigned long long* next_at(const MbfsArgs& a, uint32_t v) {
    return a.vn + size_t(v) * (2 * NW) + NW;
}

template <int NW>
__device__ __forceinline__ Bits<NW> root_mask(uint32_t k) {
    Bits<NW> m;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int cnt = int(k) - 64 * i;
        m.w[i] = cnt >= 64 ? ~0ull : (cnt <= 0 ? 0ull : ((1ull << cnt) - 1));
    }
    return m;
}

// Frontier words of level L: the history for L <= histLevels (kept so that the final
// kernel can derive every (root, vertex) level at once), else a parity pair.
template <int NW>
__device__ __forceinline__ unsigned long long* mfront(const MbfsArgs& a, uint32_t n, unsigned L) {
    if (L <= a.histLevels) return a.hist + size_t(L) * n * NW;
    return (L & 1) ? a.front1 : a.front0;
}

// The vset list of a frontier held as per-vertex words, for this workgroup's segment,
// written by the whole workgroup (block_write_vsets); each chunk's 64 word activities
// are first staged in LDS (wave w computes words 8w .. 8w+7).  Returns the count.
struct ListLds {
    unsigned tmp[32];
    unsigned long long act[kProducerBlock / 8];
};

template <int NW>
__device__ __forceinline__ unsigned segment_list(const GraphView& g, const SegView& sv,
                                                 const unsigned long long* f, uint32_t* list, ListLds& l) {
    const uint32_t ws = blockIdx.x * sv.wps, we = min(g.W, ws + sv.wps);
    constexpr unsigned kPer = kWave / 8;  // words per wave per chunk (a chunk is blockDim / 8 words)
    const unsigned wave = threadIdx.x / kWave;
    auto prep = [&](uint32_t c0) {
        bool on[kPer];
#pragma unroll
        for (unsigned j = 0; j < kPer; ++j) {  // every word's loads first (at a valid address), then the ballots
            const uint32_t w = c0 + wave * kPer + j, v = w * 64 + lane_id();
            const bool ok = w < we && v < g.n;
            Bits<NW> b;
            b.load(f, ok ? v : 0u);
            on[j] = ok && b.any();
        }
#pragma unroll
        for (unsigned j = 0; j < kPer; ++j) {
            const unsigned long long x = __ballot(on[j]);
            if (lane_id() == 0) l.act[wave * kPer + j] = x;
        }
        __syncthreads();
    };
    // chunks start at ws + k * blockDim/8, so a word's place in its chunk is (w - ws) % (blockDim/8)
    return block_write_vsets(
        g.rsetUnit, ws, we, prep, [&](uint32_t w) { return l.act[(w - ws) % (blockDim.x / 8)]; },
        list + sv.segBase[blockIdx.x], l.tmp);
}

template <int NW>
__global__ void __launch_bounds__(kProducerBlock) mbfs_init_kernel(GraphView g, MbfsArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    unsigned long long* f0 = a.hist;  // level 0
    // vertices without edges can only ever be reached by their own root: they start
    // saturated, so no slice or rset waits on them (the slots past n likewise)
    for (uint32_t v = tid; v < g.W * 64; v += nthreads) {  // whole waves: 64 consecutive v
        const bool isolated = v >= g.n || (g.csrPtr && g.csrPtr[v + 1] == g.csrPtr[v]);
        const unsigned long long bits = __ballot(isolated);
        if (lane_id() == 0) {
            a.full[v >> 6] = bits;
            a.any[v >> 6] = 0ull;
            a.any[g.W + (v >> 6)] = 0ull;
            a.dirty[v >> 6] = 0;
        }
    }
    for (size_t i = tid; i < size_t(g.n) * NW; i += nthreads) {
        f0[i] = 0ull;
        a.vn[2 * i] = 0ull;
        a.vn[2 * i + 1] = 0ull;
    }
    for (uint32_t s = tid; s < a.sv.segs; s += nthreads) {
        a.sc.news[a.sv.segs + s] = 0;
        a.sc.edges[a.sv.segs + s] = 0;
    }
    if (tid == 0) {
        level_state_reset(a.ctrl);
        a.ctrl->histTop = 0;
    }
}

// Roots set their bits and level 0, and count into their segment's slot-1 counters.
template <int NW>
__global__ void mbfs_seed_kernel(GraphView g, MbfsArgs a) {
    const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.k) return;
    const uint32_t v = a.srcs[j];
    const unsigned long long bit = 1ull << (j & 63);
    atomicOr(vis_at<NW>(a, v) + j / 64, bit);
    atomicOr(&a.hist[size_t(v) * NW + j / 64], bit);
    atomicOr(&a.any[v >> 6], 1ull << (v & 63));  // level 0: parity half 0
    a.levels[size_t(j) * g.n + v] = 0;
    const uint32_t seg = (v >> 6) / a.sv.wps;
    atomicAdd(&a.sc.news[a.sv.segs + seg], 1u);
    if (g.csrPtr) atomicAdd(&a.sc.edges[a.sv.segs + seg], (unsigned long long)(g.csrPtr[v + 1] - g.csrPtr[v]));
}

template <int NW>
__global__ void __launch_bounds__(kProducerBlock) mbfs_first_list_kernel(GraphView g, MbfsArgs a) {
    __shared__ ListLds l;
    const unsigned cnt = segment_list<NW>(g, a.sv, a.hist, a.list1, l);
    if (threadIdx.x == 0) a.sc.count[a.sv.segs + blockIdx.x] = cnt;
}

// Launched with mbfs_expand_block<NW>() threads per workgroup, one per segment: the grid
// must be resident at once (the level ends with its slowest workgroup).  64 roots fit 64
// VGPRs (8 waves per SIMD: 4 workgroups of 512 per CU); 128 roots need ~86 (5 waves per
// SIMD), so they run as workgroups of 256 (5 per CU), every workgroup resident.
template <int NW>
constexpr unsigned mbfs_expand_block() {
    return NW == 1 ? kProducerBlock : kProducerBlock / 2;
}

// Bottom-up level of a symmetric graph as a pull over the same BVSS slices.  For a
// symmetric graph a slice (row i, mask m) of rset r is also the set of edges
// i -> 8r + b (b in m), so vertex 8r + b gains OR_{slices with bit b} F[i], F = the
// previous level's frontier words.  A vset is taken by 16 lanes (two quads each per
// step, a vset holding at most 32 quads); a row's frontier words are loaded only when
// its frontier-any bit is set (the bitmap sits in L2), the 8 accumulators are OR-reduced
// over the 16 lanes, and lane b of the group ORs vertex 8r + b's new roots (those not
// yet seen there) into its `next` word: a handful of atomics per vset, none per slice,
// where the push expansion did a random record read and an atomic per slice.  Rsets
// whose 8 vertices every root has reached are skipped without reading their slices.
// The update kernel folds `next` exactly as after a push level.
template <int NW>
__device__ __forceinline__ void mbfs_pull(const GraphView& g, const MbfsArgs& a, const unsigned long long* fin,
                                          const unsigned long long* anyIn) {
    const unsigned lane = lane_id(), sub = lane & 15u, grp = lane >> 4;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const unsigned nwaves = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t v0 = wave * 4; v0 < g.V; v0 += nwaves * 4) {
        const uint32_t vs = v0 + grp;
        bool ok = vs < g.V;
        uint32_t r = 0, q0 = 0, qc = 0, open = 0;  // open: columns some root has not reached
        if (ok) {
            r = g.unitRset[vs];
            q0 = g.unitQuad[vs];
            qc = g.unitQuad[vs + 1] - q0;
            open = ~uint32_t(a.full[r >> 3] >> (8 * (r & 7))) & 0xFFu;
            ok = open != 0;
        }
        Bits<NW> acc[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[b].zero();
        for (uint32_t qq = sub; ok && qq < qc; qq += 16) {
            const uint4 r4 = stream_load_quad(g.rows, q0 + qq);
            const uint32_t m = stream_load_u32(g.masks, q0 + qq);
            const uint32_t rowv[4] = {r4.x, r4.y, r4.z, r4.w};
            bool hit[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) hit[s] = ((m >> (8 * s)) & open) != 0 && test_bit(anyIn, rowv[s]);
            Bits<NW> Fi[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                Fi[s].zero();
                if (hit[s]) Fi[s].load(fin, rowv[s]);
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint32_t ms = hit[s] ? (m >> (8 * s)) & open : 0u;
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    if ((ms >> b) & 1u)
#pragma unroll
                        for (int i = 0; i < NW; ++i) acc[b].w[i] |= Fi[s].w[i];
            }
        }
        // OR over the group's 16 lanes (xor distances below 16 stay inside the group)
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int i = 0; i < NW; ++i)
#pragma unroll
                for (int d = 8; d > 0; d >>= 1) acc[b].w[i] |= __shfl_xor(acc[b].w[i], d, kWave);
        Bits<NW> mine;
        mine.zero();
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (sub == unsigned(b)) mine = acc[b];
        const uint32_t v = 8 * r + sub;
        if (ok && sub < 8 && v < g.n && mine.any()) {
            const unsigned long long* rec = vis_at<NW>(a, v);
            unsigned long long* nr = next_at<NW>(a, v);
            bool any = false;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                const unsigned long long nv = mine.w[i] & ~(rec[i] | rec[NW + i]);
                if (nv) atomicOr(nr + i, nv);
                any |= nv != 0;
            }
            if (any) a.dirty[v >> 6] = 1;
        }
    }
}

// Top-down level as an i8 matrix product on the MFMA units (algo 1).  For a vset of
// rset r (columns 8r..8r+7) the contributions of its slices form the boolean product
// C[slice][root] = OR_b A[slice][b] AND F[8r+b][root], A = the slices' 8-bit masks,
// F = the 8 columns' frontier words.  Computed exactly as an integer product with 0/1
// i8 entries and a > 0 test: v_mfma_i32_32x32x16_i8 multiplies 32 slices (M) by 32
// roots (N) over K = 16, of which the rset's 8 columns fill K 0..7 and K 8..15 stay
// zero.  A wave takes one vset at a time (every slice of a tile shares the B operand of
// its rset) and 64 slices per step: tile 0 = slices s0..s0+31, tile 1 = s0+32..s0+63,
// each 2 x NW products (32 roots apiece).  Operand layouts (CDNA3/4, 32x32x16 i8):
// lane l holds A[l % 32][8 (l / 32) + 0..7] and B[8 (l / 32) + 0..7][l % 32] as 8 bytes;
// result register j of lane l is C[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].  A ballot
// of (C > 0) per result register therefore yields two rows' 32-root masks, and lane m
// (tile 0) or m + 32 (tile 1) collects row m's bits.  Then, as in the VALU path, a slice
// whose row is not saturated ORs the roots that have not reached the row into its
// `next` word.
typedef int mfma_i32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ long mfma_bytes_of_bits(uint32_t bits8) {
    unsigned long long v = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) v |= (unsigned long long)((bits8 >> b) & 1u) << (8 * b);
    return (long)v;
}

template <int NW>
__device__ __forceinline__ void mbfs_td_mfma(const GraphView& g, const MbfsArgs& a, const uint32_t* lin,
                                             const unsigned* pre, unsigned total,
                                             const unsigned long long* fin) {
    const unsigned lane = lane_id();
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const unsigned nwaves = (gridDim.x * blockDim.x) / kWave;
    const uint32_t* rowIds = reinterpret_cast<const uint32_t*>(g.rows);
    const uint8_t* maskBytes = reinterpret_cast<const uint8_t*>(g.masks);
    constexpr int kGroups = 2 * NW;  // 32-root groups
    // this lane's result row: tile (lane / 32), row m = lane % 32 = 8 (j / 4) + 4 h + j % 4
    const unsigned m = lane & 31u;
    const unsigned jm = 4 * (m >> 3) + (m & 3u), hm = (m >> 2) & 1u;
    for (unsigned e = wave; e < total; e += nwaves) {
        const uint32_t vs = __builtin_amdgcn_readfirstlane(list_entry(lin, a.sv, pre, e));
        const uint32_t q0 = g.unitQuad[vs];
        const uint32_t nsl = (g.unitQuad[vs + 1] - q0) * 4;
        const uint32_t r = g.unitRset[vs];
        // the rset's frontier block (wave-uniform) and the B operands built from it
        unsigned long long F[8][NW];
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int i = 0; i < NW; ++i) F[b][i] = (8 * r + b < g.n) ? fin[size_t(8 * r + b) * NW + i] : 0ull;
        long Bop[kGroups];
#pragma unroll
        for (int t = 0; t < kGroups; ++t) {
            uint32_t bits = 0;
#pragma unroll
            for (int b = 0; b < 8; ++b) bits |= uint32_t((F[b][t / 2] >> (32 * (t & 1) + (lane & 31u))) & 1ull) << b;
            Bop[t] = lane < 32 ? mfma_bytes_of_bits(bits) : 0l;
        }
        for (uint32_t s0 = 0; s0 < nsl; s0 += 64) {
            const uint32_t sl = s0 + lane;
            uint32_t row = 0, mk = 0;
            if (sl < nsl) {
                const size_t slot = size_t(q0) * 4 + sl;
                row = rowIds[slot];
                mk = maskBytes[slot];
            }
            // A operands: lanes 0..31 hold tile 0's slice masks (their own), tile 1's come
            // from lanes 32..63; lanes 32..63 hold the zero K half
            const uint32_t mkHi = __shfl(mk, (lane & 31u) + 32u, kWave);
            const long A0 = lane < 32 ? mfma_bytes_of_bits(mk) : 0l;
            const long A1 = lane < 32 ? mfma_bytes_of_bits(mkHi) : 0l;
            Bits<NW> acc;
            acc.zero();
#pragma unroll
            for (int tile = 0; tile < 2; ++tile) {
#pragma unroll
                for (int t = 0; t < kGroups; ++t) {
                    mfma_i32x16_t c = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                    c = __builtin_amdgcn_mfma_i32_32x32x16_i8(tile ? A1 : A0, Bop[t], c, 0, 0, 0);
                    unsigned long long mine = 0;
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const unsigned long long bal = __ballot(c[j] > 0);
                        mine = (unsigned(j) == jm) ? bal : mine;
                    }
                    const unsigned long long part = hm ? (mine >> 32) : (mine & 0xFFFFFFFFull);
                    if ((lane >> 5) == unsigned(tile)) acc.w[t / 2] |= part << (32 * (t & 1));
                }
            }
            bool act = mk != 0 && acc.any();
            const unsigned long long fullw = act ? a.full[row >> 6] : 0ull;
            act = act && !((fullw >> (row & 63)) & 1ull);
            if (act) {
                const unsigned long long* rec = vis_at<NW>(a, row);
                unsigned long long seen[NW];
#pragma unroll
                for (int i = 0; i < NW; ++i) seen[i] = rec[i] | rec[NW + i];
                unsigned long long* nr = next_at<NW>(a, row);
                bool any = false;
#pragma unroll
                for (int i = 0; i < NW; ++i) {
                    const unsigned long long nv = acc.w[i] & ~seen[i];
                    if (nv) atomicOr(nr + i, nv);
                    any |= nv != 0;
                }
                if (any) a.dirty[row >> 6] = 1;
            }
        }
    }
}

template <int NW, int ALGO>
__global__ void __launch_bounds__(kProducerBlock) __attribute__((amdgpu_waves_per_eu(NW == 1 ? 8 : 5)))
mbfs_expand_kernel(GraphView g, MbfsArgs a, unsigned Lk) {
    const unsigned L = level_of(a.ctrl, a.levelFromCtrl, Lk);
    __shared__ unsigned pre[kMaxSegs];
    __shared__ unsigned long long tmp64[16];
    Ctrl* c = a.ctrl;
    const uint32_t segs = a.sv.segs;
    DirectionParams dp;
    dp.direction = a.direction;
    dp.canBottomUp = g.csrPtr != nullptr || a.pull;
    dp.alpha = a.alpha;
    dp.beta = a.beta;
    dp.totalEdges = (unsigned long long)a.k * g.nnz;
    dp.totalItems = double(g.n) * a.k;
    bool bottomUp = false;
    unsigned total = 0;  // list entries of the level (top-down)
    if (!level_begin(c, a.sc, segs, L, dp, bottomUp, pre, &total)) return;
    const unsigned lane = lane_id();
    const unsigned long long* fin = mfront<NW>(a, g.n, L - 1);
    unsigned long long* fout = mfront<NW>(a, g.n, L);
    const bool writeLevels = L > a.histLevels;
    const unsigned sOut = (L + 1) % 3;
    if (!bottomUp && ALGO == 1) {
        mbfs_td_mfma<NW>(g, a, (L & 1) ? a.list1 : a.list0, pre, total, fin);
        return;
    }
    if (bottomUp && a.pull) {
        mbfs_pull<NW>(g, a, fin, a.any + size_t((L - 1) & 1) * g.W);
        return;
    }
    if (!bottomUp) {
        const uint32_t* lin = (L & 1) ? a.list1 : a.list0;

        const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
        const unsigned nwaves = (gridDim.x * blockDim.x) / kWave;
        // m list entries per wave step (a short list is spread over as many waves)
        const unsigned m = min(unsigned(kWave), max(1u, (total + nwaves - 1) / nwaves));
        for (unsigned base = wave * m; base < total; base += nwaves * m) {
            const unsigned e = base + lane;
            uint32_t q0 = 0, qc = 0, r = 0;
            if (lane < m && e < total) {
                const uint32_t vs = list_entry(lin, a.sv, pre, e);
                q0 = g.unitQuad[vs];
                qc = g.unitQuad[vs + 1] - q0;
                r = g.unitRset[vs];
            }
            const unsigned incl = wave_inclusive_scan(qc);
            const unsigned excl = incl - qc;
            const unsigned tq = __shfl(incl, kWave - 1, kWave);
            for (unsigned t0 = 0; t0 < tq; t0 += kWave) {
                const unsigned t = t0 + lane;
                const unsigned owner = wave_owner(excl, min(t, tq - 1));
                const uint32_t oq0 = __shfl(q0, owner, kWave);
                const unsigned oex = __shfl(excl, owner, kWave);
                const uint32_t orr = __shfl(r, owner, kWave);
                if (t < tq) {
                    const uint32_t q = oq0 + (t - oex);
                    const uint4 r4 = stream_load_quad(g.rows, q);
                    const uint32_t m = stream_load_u32(g.masks, q);
                    const uint32_t um = (m | (m >> 8) | (m >> 16) | (m >> 24)) & 0xFF;
                    Bits<NW> F[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        F[b].zero();
                        if ((um >> b) & 1) F[b].load(fin, orr * 8 + b);
                    }
                    const uint32_t rowv[4] = {r4.x, r4.y, r4.z, r4.w};
                    // Phased so that the four slices' loads overlap: the roots arriving
                    // per slice (registers), then the saturated-row bits (L2), then the
                    // row records, and only then the atomics (an atomic on a record may
                    // alias a later slice's load, so mixing them serialises the slices).
                    Bits<NW> acc[4];
                    bool act[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const uint32_t ms = (m >> (8 * s)) & 0xFF;
                        acc[s].zero();
#pragma unroll
                        for (int b = 0; b < 8; ++b)
                            if ((ms >> b) & 1)
#pragma unroll
                                for (int i = 0; i < NW; ++i) acc[s].w[i] |= F[b].w[i];
                        act[s] = acc[s].any();
                    }
                    unsigned long long fullw[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) fullw[s] = act[s] ? a.full[(act[s] ? rowv[s] : 0u) >> 6] : 0ull;
#pragma unroll
                    for (int s = 0; s < 4; ++s) act[s] = act[s] && !((fullw[s] >> (rowv[s] & 63)) & 1ull);
                    unsigned long long seen[4][NW];
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const unsigned long long* rec = vis_at<NW>(a, act[s] ? rowv[s] : 0u);
#pragma unroll
                        for (int i = 0; i < NW; ++i) seen[s][i] = act[s] ? (rec[i] | rec[NW + i]) : ~0ull;
                    }
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        unsigned long long* nr = next_at<NW>(a, act[s] ? rowv[s] : 0u);
                        bool any = false;
#pragma unroll
                        for (int i = 0; i < NW; ++i) {
                            const unsigned long long nv = acc[s].w[i] & ~seen[s][i];
                            if (act[s] && nv) atomicOr(nr + i, nv);
                            any |= act[s] && nv != 0;
                        }
                        if (any) a.dirty[rowv[s] >> 6] = 1;  // a byte store: idempotent
                    }
                }
            }
        }
        return;
    }
    // bottom-up over this workgroup's segment: wave per word, lane per vertex.  A vertex
    // of degree > kHeavyDeg is deferred to a second pass in which a whole wave scans its
    // neighbours 64 at a time: with many roots a scan rarely stops early, and one lane
    // walking a hub's list would hold its wave (and the level) for the whole list.
    constexpr uint32_t kHeavyDeg = 64;
    constexpr unsigned kHeavyCap = 2048;
    __shared__ uint32_t heavyQ[kHeavyCap];
    __shared__ unsigned heavyN;
    if (threadIdx.x == 0) heavyN = 0;
    __syncthreads();
    const Bits<NW> rootMask = root_mask<NW>(a.k);
    const uint32_t ws = blockIdx.x * a.sv.wps, we = min(g.W, ws + a.sv.wps);
    const unsigned wib = threadIdx.x / kWave, wpb = blockDim.x / kWave;
    unsigned long long news = 0, edges = 0;
    for (uint32_t w = ws + wib; w < we; w += wpb) {
        const uint32_t v = w * 64 + lane;
        if (v >= g.n) continue;
        Bits<NW> vis, U, nw;
        vis.loadp(vis_at<NW>(a, v));
        nw.zero();
        bool need = false;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            U.w[i] = ~vis.w[i] & rootMask.w[i];
            need |= U.w[i] != 0;
        }
        const uint32_t p0 = need ? g.csrPtr[v] : 0u, p1 = need ? g.csrPtr[v + 1] : 0u;
        if (need && p1 - p0 > kHeavyDeg) {
            const unsigned slot = atomicAdd(&heavyN, 1u);
            if (slot < kHeavyCap) {
                heavyQ[slot] = v;
                need = false;  // its frontier word is written by the heavy pass
            }
        }
        if (need) {
            Bits<NW> acc;
            acc.zero();
            for (uint32_t p = p0; p < p1; ++p) {
                const uint32_t u = g.csrIdx[p];
                bool all = true;
#pragma unroll
                for (int i = 0; i < NW; ++i) {
                    acc.w[i] |= fin[size_t(u) * NW + i];
                    all &= (acc.w[i] & U.w[i]) == U.w[i];
                }
                if (all) break;
            }
#pragma unroll
            for (int i = 0; i < NW; ++i) nw.w[i] = acc.w[i] & U.w[i];
            if (nw.any()) {
#pragma unroll
                for (int i = 0; i < NW; ++i) vis.w[i] |= nw.w[i];
                vis.storep(vis_at<NW>(a, v));
                if (writeLevels)
                    for (int i = 0; i < NW; ++i)
                        for (unsigned long long b = nw.w[i]; b; b &= b - 1)
                            a.levels[size_t(64 * i + __builtin_ctzll(b)) * g.n + v] = int32_t(L);
                news += nw.popc();
                edges += (unsigned long long)(p1 - p0) * nw.popc();
            }
        }
        nw.store(fout, v);
    }
    __syncthreads();
    const unsigned nh = min(heavyN, kHeavyCap);
    for (unsigned h = wib; h < nh; h += wpb) {
        const uint32_t v = heavyQ[h];
        Bits<NW> vis, U, acc;
        vis.loadp(vis_at<NW>(a, v));
#pragma unroll
        for (int i = 0; i < NW; ++i) U.w[i] = ~vis.w[i] & rootMask.w[i];
        acc.zero();
        const uint32_t p0 = g.csrPtr[v], p1 = g.csrPtr[v + 1];
        for (uint32_t pb = p0; pb < p1; pb += kWave) {
            const uint32_t p = pb + lane;
            if (p < p1) {
                const uint32_t u = g.csrIdx[p];
#pragma unroll
                for (int i = 0; i < NW; ++i) acc.w[i] |= fin[size_t(u) * NW + i];
            }
            bool all = true;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                for (int d = kWave / 2; d > 0; d >>= 1) acc.w[i] |= __shfl_xor(acc.w[i], d, kWave);
                all &= (acc.w[i] & U.w[i]) == U.w[i];
            }
            if (all) break;  // acc is the same on every lane: uniform exit
        }
        if (lane == 0) {
            Bits<NW> nw;
#pragma unroll
            for (int i = 0; i < NW; ++i) nw.w[i] = acc.w[i] & U.w[i];
            if (nw.any()) {
#pragma unroll
                for (int i = 0; i < NW; ++i) vis.w[i] |= nw.w[i];
                vis.storep(vis_at<NW>(a, v));
                if (writeLevels)
                    for (int i = 0; i < NW; ++i)
                        for (unsigned long long b = nw.w[i]; b; b &= b - 1)
                            a.levels[size_t(64 * i + __builtin_ctzll(b)) * g.n + v] = int32_t(L);
                news += nw.popc();
                edges += (unsigned long long)(p1 - p0) * nw.popc();
            }
            nw.store(fout, v);
        }
    }
    news = block_sum_u64(news, tmp64);
    edges = block_sum_u64(edges, tmp64);
    if (threadIdx.x == 0) {
        a.sc.news[size_t(sOut) * segs + blockIdx.x] = unsigned(min(news, 0xFFFFFFFFull));
        a.sc.edges[size_t(sOut) * segs + blockIdx.x] = edges;
    }
}

template <int NW>
__global__ void __launch_bounds__(kProducerBlock) BLEST_RESIDENT_GRID mbfs_update_kernel(GraphView g, MbfsArgs a, unsigned Lk) {
    const unsigned L = level_of(a.ctrl, a.levelFromCtrl, Lk);
    __shared__ ListLds l;
    __shared__ unsigned long long tmp64[16];
    Ctrl* c = a.ctrl;
    if (ld_relaxed(&c->done)) return;
    const bool bottomUp = (c->state[L & 1].dir & 1u) != 0;
    unsigned long long* fout = mfront<NW>(a, g.n, L);
    const bool writeLevels = L > a.histLevels;
    uint32_t* lout = (L & 1) ? a.list0 : a.list1;
    const uint32_t segs = a.sv.segs;
    const unsigned sOut = (L + 1) % 3;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->histTop = L;  // read only after this kernel
    if (bottomUp && !a.pull) {
        const unsigned cnt = segment_list<NW>(g, a.sv, fout, lout, l);
        if (threadIdx.x == 0) a.sc.count[size_t(sOut) * segs + blockIdx.x] = cnt;
        return;
    }
    const uint32_t ws = blockIdx.x * a.sv.wps, we = min(g.W, ws + a.sv.wps);
    uint32_t first, last;
    wave_words(ws, we, threadIdx.x / kWave, blockDim.x / kWave, first, last);
    unsigned long long news = 0, edges = 0;
    const Bits<NW> rootMask = root_mask<NW>(a.k);
    // groups of kGroup words: the group's records (visited and next, one line per vertex)
    // are all loaded before any is used, so a group costs one memory round trip
    constexpr int kGroup = NW == 1 ? 4 : 2;  // (4 at 128 roots spills past the 64 VGPRs of a resident grid)
    // A word without a dirty mark holds no vertex whose `next` word was set this level:
    // its records are not read, its frontier words are stored as zeros.
    for (uint32_t w0 = first; w0 < last; w0 += kGroup) {
        Bits<NW> nx[kGroup], vs[kGroup];
        bool ok[kGroup], dj[kGroup];
#pragma unroll
        for (int j = 0; j < kGroup; ++j) dj[j] = w0 + j < last && a.dirty[w0 + j] != 0;
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const uint32_t v = (w0 + j) * 64 + lane_id();
            ok[j] = w0 + j < last && v < g.n;
            nx[j].zero();
            vs[j].zero();
            if (ok[j] && dj[j]) {
                nx[j].loadp(next_at<NW>(a, v));
                vs[j].loadp(vis_at<NW>(a, v));
            }
        }
        uint32_t deg[kGroup];
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const uint32_t v = (w0 + j) * 64 + lane_id();
            deg[j] = (ok[j] && g.csrPtr && nx[j].any()) ? g.csrPtr[v + 1] - g.csrPtr[v] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const uint32_t w = w0 + j, v = w * 64 + lane_id();
            bool nowFull = false;
            if (ok[j]) {
                const Bits<NW>& nw = nx[j];
                if (nw.any()) {
                    Bits<NW> vis = vs[j], z;
                    z.zero();
                    z.storep(next_at<NW>(a, v));
                    nowFull = true;
#pragma unroll
                    for (int i = 0; i < NW; ++i) {
                        vis.w[i] |= nw.w[i];
                        nowFull &= vis.w[i] == rootMask.w[i];
                    }
                    vis.storep(vis_at<NW>(a, v));
                    const unsigned pc = nw.popc();
                    if (writeLevels)
                        for (int i = 0; i < NW; ++i)
                            for (unsigned long long b = nw.w[i]; b; b &= b - 1)
                                a.levels[size_t(64 * i + __builtin_ctzll(b)) * g.n + v] = int32_t(L);
                    news += pc;
                    edges += (unsigned long long)deg[j] * pc;
                }
                nw.store(fout, v);
            }
            const unsigned long long fullBits = __ballot(nowFull);
            const unsigned long long anyBits = __ballot(ok[j] && nx[j].any());
            if (lane_id() == 0 && w < last) {  // this wave owns word w here
                if (fullBits) a.full[w] |= fullBits;
                a.any[size_t(L & 1) * g.W + w] = anyBits;
                if (dj[j]) a.dirty[w] = 0;
            }
        }
    }
    // the counts first (their barriers also publish the activity words to the other waves),
    // so that nothing of the word pass stays live through the list writing
    news = block_sum_u64(news, tmp64);
    edges = block_sum_u64(edges, tmp64);
    if (threadIdx.x == 0) {
        a.sc.news[size_t(sOut) * segs + blockIdx.x] = unsigned(min(news, 0xFFFFFFFFull));
        a.sc.edges[size_t(sOut) * segs + blockIdx.x] = edges;
    }
    const unsigned long long* anyL = a.any + size_t(L & 1) * g.W;
    const uint32_t ws2 = blockIdx.x * a.sv.wps, we2 = min(g.W, ws2 + a.sv.wps);
    const unsigned running = block_write_vsets(
        g.rsetUnit, ws2, we2, [](uint32_t) {}, [&](uint32_t w) { return anyL[w]; }, lout + a.sv.segBase[blockIdx.x],
        l.tmp);
    if (threadIdx.x == 0) a.sc.count[size_t(sOut) * segs + blockIdx.x] = running;
}

// Every (root, vertex) level at once, thread per vertex.  The level of root j is the
// history level L whose frontier word holds bit j; with L <= 16 its five bits are read
// from five bit-planes (plane b = OR of the words of the levels with bit b set), so a
// vertex costs its history words in and its k levels out, nothing else.  Pairs found
// past the history were written when found; the rest are unreached (-1).  Root-major
// rows: for each root the threads store consecutive levels.  Idempotent: enqueued ahead
// of each completion poll, it reads only levels whose words are complete.
constexpr unsigned kFinalBlock = 256;

template <int NW>
__global__ void __launch_bounds__(kFinalBlock) mbfs_final_kernel(GraphView g, MbfsArgs a) {
    const unsigned top = min(__hip_atomic_load(&a.ctrl->histTop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                             a.histLevels);
    for (uint32_t v = blockIdx.x * kFinalBlock + threadIdx.x; v < g.n; v += gridDim.x * kFinalBlock) {
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            if (64 * i >= int(a.k)) break;
            unsigned long long plane[5] = {0, 0, 0, 0, 0}, inHist = 0;
            for (unsigned L = 0; L <= top; ++L) {
                const unsigned long long x = a.hist[(size_t(L) * g.n + v) * NW + i];
                inHist |= x;
#pragma unroll
                for (int b = 0; b < 5; ++b)
                    if ((L >> b) & 1u) plane[b] |= x;
            }
            const unsigned long long vis = vis_at<NW>(a, v)[i];
            const unsigned jn = min(64u, a.k - 64u * i);
            for (unsigned j = 0; j < jn; ++j) {
                int32_t* out = a.levels + size_t(64 * i + j) * g.n + v;
                // streaming stores: the k x n output is written once and not read here
                if ((inHist >> j) & 1ull) {
                    int32_t l = 0;
#pragma unroll
                    for (int b = 0; b < 5; ++b) l |= int32_t((plane[b] >> j) & 1ull) << b;
                    __builtin_nontemporal_store(l, out);
                } else if (!((vis >> j) & 1ull)) {
                    __builtin_nontemporal_store(int32_t(-1), out);
                }
            }
        }
    }
}

__global__ void unpermute_rows_kernel(int32_t* out, const int32_t* in, const uint32_t* oldToNew, uint32_t n,
                                      uint32_t rows) {
    for (uint32_t r = blockIdx.y; r < rows; r += gridDim.y)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            out[size_t(r) * n + i] = in[size_t(r) * n + oldToNew[i]];
}

// Per-root totals of a multi-source run (rows of k x n levels, -1 = unreached).
struct RootStats {
    unsigned levels;  // max level + 1
    unsigned visited;
    unsigned long long degreeSum;
};

__global__ void root_stats_clear_kernel(RootStats* st, uint32_t k) {
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) st[i] = RootStats{0u, 0u, 0ull};
}

__global__ void __launch_bounds__(256) root_stats_kernel(RootStats* st, const int32_t* levels, const uint32_t* csrPtr,
                                                         uint32_t n) {
    __shared__ unsigned sMax[4], sCnt[4];
    __shared__ unsigned long long sDeg[4];
    const uint32_t r = blockIdx.y;
    const int32_t* row = levels + size_t(r) * n;
    unsigned mx = 0, cnt = 0;
    unsigned long long deg = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t l = row[i];
        if (l >= 0) {
            mx = max(mx, unsigned(l) + 1);
            ++cnt;
            if (csrPtr) deg += csrPtr[i + 1] - csrPtr[i];
        }
    }
    for (int d = kWave / 2; d > 0; d >>= 1) {
        mx = max(mx, unsigned(__shfl_xor(int(mx), d, kWave)));
        cnt += __shfl_xor(cnt, d, kWave);
        deg += __shfl_xor(deg, d, kWave);
    }
    const unsigned w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        sMax[w] = mx;
        sCnt[w] = cnt;
        sDeg[w] = deg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned i = 1; i < blockDim.x / kWave; ++i) {
            mx = max(mx, sMax[i]);
            cnt += sCnt[i];
            deg += sDeg[i];
        }
        if (mx) atomicMax(&st[r].levels, mx);
        if (cnt) atomicAdd(&st[r].visited, cnt);
        if (deg) atomicAdd(&st[r].degreeSum, deg);
    }
}

}  // namespace dev

static void mbfs_ensure_state(Query& q, uint32_t nw) {
    // (declared in dgraph.hpp: the history depth also sizes the query footprint)
    DeviceGraph& g = *q.g;
    if (q.mbfsWords >= nw) return;
    for (void* p : {static_cast<void*>(q.mVis), static_cast<void*>(q.mFront0), static_cast<void*>(q.mFront1),
                    static_cast<void*>(q.mHist), static_cast<void*>(q.mFull), static_cast<void*>(q.mAny),
                    static_cast<void*>(q.mDirty)})
        q.release(p);
    // a new layout invalidates every captured multi-source level chunk
    for (int k : {kGraphMbfs1, kGraphMbfs2})
        if (q.graphs[k].exec) {
            BLEST_HIP(hipGraphExecDestroy(q.graphs[k].exec));
            q.graphs[k].exec = nullptr;
        }
    const size_t words = size_t(g.n) * nw;
    q.mVis = q.alloc<unsigned long long>(2 * words);  // visited + next, interleaved per vertex
    q.mFront0 = q.alloc<unsigned long long>(words);
    q.mFront1 = q.alloc<unsigned long long>(words);
    q.mHistLevels = mbfs_history_levels(g.n, nw);
    q.mHist = q.alloc<unsigned long long>(words * (q.mHistLevels + 1));
    q.mFull = q.alloc<unsigned long long>(g.W);
    q.mAny = q.alloc<unsigned long long>(2 * size_t(g.W));
    q.mDirty = q.alloc<uint8_t>(g.W);
    if (!q.srcs) q.srcs = q.alloc<uint32_t>(128);
    if (!q.rootStats) q.rootStats = q.alloc<dev::RootStats>(128);
    q.mbfsWords = nw;
}

template <int NW, int ALGO>
static void mbfs_enqueue_levels(Query& q, const dev::MbfsArgs& a) {
    DeviceGraph& g = *q.g;
    const GraphView gv = g.view();
    const dim3 grid = level_grid(g), block(dev::kProducerBlock), eblock(dev::mbfs_expand_block<NW>());
    hipLaunchKernelGGL(dev::mbfs_init_kernel<NW>, grid, block, 0, q.stream, gv, a);
    hipLaunchKernelGGL(dev::mbfs_seed_kernel<NW>, dim3(2), dim3(64), 0, q.stream, gv, a);
    hipLaunchKernelGGL(dev::mbfs_first_list_kernel<NW>, grid, block, 0, q.stream, gv, a);
    dev::MbfsArgs ar = a;
    ar.levelFromCtrl = 1;
    run_levels(
        q,
        [&](unsigned L) {
            hipLaunchKernelGGL((dev::mbfs_expand_kernel<NW, ALGO>), grid, eblock, 0, q.stream, gv, a, L);
            hipLaunchKernelGGL(dev::mbfs_update_kernel<NW>, grid, block, 0, q.stream, gv, a, L);
        },
        [&] {
            const uint32_t fb = std::min<uint32_t>((g.n + dev::kFinalBlock - 1) / dev::kFinalBlock, uint32_t(g.cus) * 8);
            hipLaunchKernelGGL(dev::mbfs_final_kernel<NW>, dim3(fb), dim3(dev::kFinalBlock), 0, q.stream, gv, a);
        },
        &q.graphs[NW == 1 ? kGraphMbfs1 : kGraphMbfs2],
        [&](unsigned k) {
            hipLaunchKernelGGL((dev::mbfs_expand_kernel<NW, ALGO>), grid, eblock, 0, q.stream, gv, ar, k);
            hipLaunchKernelGGL(dev::mbfs_update_kernel<NW>, grid, block, 0, q.stream, gv, ar, k);
        },
        launch_key(a.alpha, a.beta, a.direction, a.k | (uint32_t(ALGO) << 16)));
    BLEST_HIP(hipGetLastError());
}

// hostLevels: k x n contiguous rows, or rows: k separate arrays of n (either may be null).
static void mbfs_run(Query& q, const uint32_t* srcs, uint32_t k, int32_t* hostLevels, int32_t* const* rows,
                     const blest_mbfs_opts* o, blest_bfs_stats* st, blest_bfs_stats* perRoot = nullptr) {
    DeviceGraph& g = *q.g;
    if (k == 0 || k > 128) throw Error(BLEST_ERR_INVALID_ARGUMENT, "multi-source batches hold 1..128 roots");
    if (k > q.maxSources) throw Error(BLEST_ERR_INVALID_ARGUMENT, "more roots than the query's max_sources");
    const int algo = o ? o->algo : 0;
    if (algo != 0 && algo != 1) throw Error(BLEST_ERR_INVALID_ARGUMENT, "algo must be 0 (bit-parallel) or 1 (i8 MFMA)");
    const uint32_t nw = (k > 64 || (o && o->batch > 64)) ? 2 : 1;
    // result destinations are checked before anything runs
    if (hostLevels) check_result_range(hostLevels, sizeof(int32_t) * size_t(g.n) * k, g.device, "levels");
    if (rows)
        for (uint32_t j = 0; j < k; ++j) {
            if (!rows[j]) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null result row");
            check_result_range(rows[j], sizeof(int32_t) * size_t(g.n), g.device, "result row");
        }
    mbfs_ensure_state(q, nw);
    std::vector<uint32_t> mapped(k);
    for (uint32_t j = 0; j < k; ++j) mapped[j] = map_source(g, srcs[j]);
    BLEST_HIP(hipMemcpyAsync(q.srcs, mapped.data(), 4 * size_t(k), hipMemcpyHostToDevice, q.stream));
    dev::MbfsArgs a;
    a.levels = q.levels;
    a.vn = q.mVis;
    a.front0 = q.mFront0;
    a.front1 = q.mFront1;
    a.hist = q.mHist;
    a.histLevels = q.mHistLevels;
    a.full = q.mFull;
    a.any = q.mAny;
    a.dirty = q.mDirty;
    a.pull = g.symmetric ? 1 : 0;
    a.list0 = q.list0;
    a.list1 = q.list1;
    a.srcs = q.srcs;
    a.k = k;
    a.ctrl = q.ctrl;
    a.sv = g.seg_view();
    a.sc = q.seg_counters();
    // Direction.  A symmetric graph pulls over the BVSS slices on bottom-up levels
    // (mbfs_pull), and switches by the graph's thresholds (direction 0) unless one
    // direction is asked for.  Otherwise top-down unless bottom-up (the CSR scan) is
    // asked for explicitly (direction 2): with 64-128 roots a per-vertex scan must find
    // every root still missing at a vertex, so it seldom stops early; on R-MAT scale 22
    // (128 roots) that switching schedule took 21 ms per batch and top-down only 9.6 ms.
    if (g.symmetric)
        a.direction = g.opts.direction;
    else
        a.direction = (g.hasCsr && g.opts.direction == 2) ? 2 : 1;
    a.alpha = g.opts.alpha;
    a.beta = g.opts.beta;
    a.levelFromCtrl = 0;
    BLEST_HIP(hipEventRecord(q.ev0, q.stream));
    if (nw == 1) {
        if (algo == 1) mbfs_enqueue_levels<1, 1>(q, a);
        else mbfs_enqueue_levels<1, 0>(q, a);
    } else {
        if (algo == 1) mbfs_enqueue_levels<2, 1>(q, a);
        else mbfs_enqueue_levels<2, 0>(q, a);
    }
    BLEST_HIP(hipEventRecord(q.ev1, q.stream));
    BLEST_HIP(hipStreamSynchronize(q.stream));
    dev::Ctrl c;
    BLEST_HIP(hipMemcpyAsync(&c, q.ctrl, sizeof(c), hipMemcpyDeviceToHost, q.stream));
    BLEST_HIP(hipStreamSynchronize(q.stream));
    q.levelHint = c.levels;
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
    if (perRoot) {
        auto* rsDev = static_cast<dev::RootStats*>(q.rootStats);
        hipLaunchKernelGGL(dev::root_stats_clear_kernel, dim3(1), dim3(128), 0, q.stream, rsDev, k);
        const uint32_t bx = std::min<uint32_t>(256, (g.n + 255) / 256);
        hipLaunchKernelGGL(dev::root_stats_kernel, dim3(bx, k), dim3(256), 0, q.stream, rsDev, q.levels,
                           g.hasCsr ? g.csrPtr : nullptr, g.n);
        BLEST_HIP(hipGetLastError());
        std::vector<dev::RootStats> rs(k);
        BLEST_HIP(hipMemcpyAsync(rs.data(), rsDev, sizeof(dev::RootStats) * k, hipMemcpyDeviceToHost, q.stream));
        BLEST_HIP(hipStreamSynchronize(q.stream));
        for (uint32_t j = 0; j < k; ++j) {
            perRoot[j] = st ? *st : blest_bfs_stats{};
            perRoot[j].kernel_ms = ms;
            perRoot[j].total_levels = rs[j].levels;
            perRoot[j].no_visited = rs[j].visited;
            perRoot[j].edges_traversed_component = rs[j].degreeSum / 2;
        }
    }
    if (hostLevels || rows) {
        const int32_t* src = q.levels;
        if (g.hasPerm) {
            hipLaunchKernelGGL(dev::unpermute_rows_kernel, dim3(256, std::min<uint32_t>(k, 65535)), dim3(256), 0,
                               q.stream, q.levelsOut, q.levels, g.permOldToNew, g.n, k);
            BLEST_HIP(hipGetLastError());
            src = q.levelsOut;
        }
        if (hostLevels)
            BLEST_HIP(hipMemcpyAsync(hostLevels, src, sizeof(int32_t) * size_t(g.n) * k, hipMemcpyDefault,
                                     q.stream));
        else
            for (uint32_t j = 0; j < k; ++j)
                BLEST_HIP(hipMemcpyAsync(rows[j], src + size_t(g.n) * j, sizeof(int32_t) * size_t(g.n),
                                         hipMemcpyDefault, q.stream));
        BLEST_HIP(hipStreamSynchronize(q.stream));
    }
}

}  // namespace blest

using namespace blest;

extern "C" blest_status blest_mbfs(blest_query* qh, const uint32_t* srcs, uint32_t k, int32_t* levels,
                                   const blest_mbfs_opts* o, blest_bfs_stats* st) {
    return guarded([&] {
        if (!qh || !srcs) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        mbfs_run(q, srcs, k, levels, nullptr, o, st);
    });
}

blest_status blest::mbfs_rows(blest_query* qh, const uint32_t* srcs, uint32_t k, int32_t* const* rows,
                              blest_bfs_stats* st, blest_bfs_stats* perRoot) {
    return guarded([&] {
        if (!qh || !srcs || !rows) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        mbfs_run(q, srcs, k, nullptr, rows, nullptr, st, perRoot);
    });
}
