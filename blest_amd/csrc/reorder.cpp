// Vertex reorderings (host).  Results of BFS and CC do not depend on the ordering;
// the ordering only shapes the BVSS slices (how many rows share a slice) and with
// that the bytes the traversal streams.
#include "reorder.hpp"

#include <cmath>
#include <numeric>

namespace blest {

std::vector<uint32_t> order_natural(const CSC& g) {
    std::vector<uint32_t> inv(g.n);
    std::iota(inv.begin(), inv.end(), 0u);
    return inv;
}

// CSC::random (CSC.cuh:535-554), seeded instead of std::random_device.
std::vector<uint32_t> order_random(const CSC& g, uint64_t seed) {
    std::vector<uint32_t> perm = seeded_permutation(g.n, seed, 11);
    std::vector<uint32_t> inv(g.n);
    for (uint32_t i = 0; i < g.n; ++i) inv[perm[i]] = i;
    return inv;
}

// CSC::degreeSort (CSC.cuh:968-992): by degree, ties by vertex ID.
std::vector<uint32_t> order_degree(const CSC& g, bool ascending) {
    std::vector<uint32_t> ids(g.n);
    std::iota(ids.begin(), ids.end(), 0u);
    std::sort(ids.begin(), ids.end(), [&](uint32_t x, uint32_t y) {
        uint32_t dx = g.degree(x), dy = g.degree(y);
        if (dx != dy) return ascending ? dx < dy : dx > dy;
        return x < y;
    });
    std::vector<uint32_t> inv(g.n);
    for (uint32_t k = 0; k < g.n; ++k) inv[ids[k]] = k;
    return inv;
}

// ------------------------------------------------------------------------- RCM
namespace {

// findPseudoPeripheralNode (CSC.cuh:994-1057): repeated BFS from a random vertex of
// the previous BFS's last level until the depth has not changed for 10 rounds.  The
// reference draws the restart with glibc rand(); here it is a seeded counter draw.
uint32_t pseudo_peripheral(const CSC& s, uint32_t start, std::vector<uint32_t>& mark,
                           uint32_t& stamp, std::vector<uint32_t>& q, uint64_t key,
                           uint64_t& draw) {
    uint32_t maxLevel = 0, trial = 0;
    while (trial < 10) {
        ++stamp;
        uint32_t qs = 0, qe = 0;
        q[qe++] = start;
        mark[start] = stamp;
        uint32_t levelStart = 0, levelEnd = 1, levels = 1;
        while (qs < qe) {
            uint32_t u = q[qs++];
            for (uint32_t p = s.colPtrs[u]; p < s.colPtrs[u + 1]; ++p) {
                uint32_t v = s.rows[p];
                if (mark[v] != stamp) {
                    mark[v] = stamp;
                    q[qe++] = v;
                }
            }
            if (qs == levelEnd && qs < qe) {
                levelStart = qs;
                levelEnd = qe;
                ++levels;
            }
        }
        trial = (levels == maxLevel) ? trial + 1 : 0;
        maxLevel = levels;
        start = q[levelStart + uint32_t(crand(key, draw++) % (levelEnd - levelStart))];
    }
    return start;
}

}  // namespace

// CSC::rcm (CSC.cuh:1059-1142): per component, BFS from a pseudo-peripheral vertex,
// neighbours visited in ascending degree (stable here; std::sort in the reference),
// numbered in BFS order (the reference does not reverse it despite the name).
std::vector<uint32_t> order_rcm(const CSC& g, uint64_t seed) {
    CSC sym;
    const CSC* s = &g;
    if (!csc_is_symmetric(g)) {
        sym = csc_symmetrize(g);
        s = &sym;
    }
    const uint32_t n = g.n;
    std::vector<uint32_t> inv(n), q(std::max<uint32_t>(n, 1)), mark(n, 0), pq(std::max<uint32_t>(n, 1)),
        nbrs;
    std::vector<uint8_t> visited(n, 0);
    uint32_t stamp = 0, permuted = 0;
    uint64_t key = stream_key(seed, 12), draw = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (visited[i]) continue;
        uint32_t start = pseudo_peripheral(*s, i, mark, stamp, pq, key, draw);
        uint32_t qs = 0, qe = 0;
        q[qe++] = start;
        visited[start] = 1;
        while (qs < qe) {
            uint32_t u = q[qs++];
            inv[u] = permuted++;
            nbrs.clear();
            for (uint32_t p = s->colPtrs[u]; p < s->colPtrs[u + 1]; ++p) {
                uint32_t v = s->rows[p];
                if (!visited[v]) {
                    visited[v] = 1;
                    nbrs.push_back(v);
                }
            }
            std::stable_sort(nbrs.begin(), nbrs.end(),
                             [&](uint32_t x, uint32_t y) { return s->degree(x) < s->degree(y); });
            for (uint32_t v : nbrs) q[qe++] = v;
        }
    }
    return inv;
}

// -------------------------------------------------------------- Jaccard window
// jackardWithWindow (CSC.cuh:717-920).  Columns are grouped into windows of `window`
// consecutive IDs that are ordered independently (and in parallel); a column never
// leaves its window.  Within a window, each run of sliceSize new positions (one
// future real slice set) starts from the unplaced column of largest degree (ties:
// largest ID) and is filled greedily with the candidate of largest Jaccard similarity
// inter / (union + deg - inter) to the union of rows already in the run (first best
// wins ties); candidates are unplaced window columns sharing a row with the run.
// When no candidate is left the next singleton is taken.
std::vector<uint32_t> order_jaccard_window(const CSC& g, uint32_t sliceSize, uint32_t window,
                                           unsigned threads) {
    if (sliceSize == 0 || window == 0 || window % sliceSize != 0)
        throw Error(BLEST_ERR_INVALID_ARGUMENT, "Jaccard window must be a positive multiple of the slice size");
    threads = hw_threads(threads);
    const uint32_t n = g.n;
    CSC tr;
    const CSC* t = &g;  // rows -> columns index: the transpose (g itself when symmetric)
    if (!csc_is_symmetric(g, threads)) {
        tr = csc_transpose(g, threads);
        t = &tr;
    }
    std::vector<uint32_t> inv(n);
    const uint32_t noWindows = (n + window - 1) / window;
    struct Scratch {
        std::vector<uint8_t> placed;     // per window column
        std::vector<uint32_t> inter;     // per window column
        std::vector<uint32_t> queue;     // candidate columns
        std::vector<uint32_t> rowMark;   // per row: epoch of the run that holds it
        std::vector<uint32_t> rowStart;  // per row: first transpose position >= windowStart
        std::vector<uint32_t> rowWin;    // per row: window for which rowStart is valid (+1)
        std::vector<uint32_t> bySize;    // window columns by (degree desc, id desc)
        std::vector<uint32_t> rowLen;    // per row: kPooled | length of its run in `pool`
        std::vector<uint32_t> pool;      // long rows' window columns, unplaced ones first
        uint32_t epoch = 0;
    };
    // Rows with at least kLongRun columns in a window get a private copy of that run,
    // compacted whenever more than half of it has been placed (ascending order kept, so
    // the candidate queue fills in the same order).
    constexpr uint32_t kLongRun = 64, kPooled = 0x80000000u;
    std::vector<Scratch> scratch(threads);
    parallel_dynamic(noWindows, threads, 1, [&](uint64_t w, unsigned tid) {
        Scratch& S = scratch[tid];
        const uint32_t ws = uint32_t(w) * window, we = std::min<uint64_t>(uint64_t(ws) + window, n);
        const uint32_t len = we - ws;
        if (S.rowMark.empty()) {
            S.rowMark.assign(n, 0);
            S.rowStart.assign(n, 0);
            S.rowWin.assign(n, 0);
            S.rowLen.assign(n, 0);
        }
        S.pool.clear();
        S.placed.assign(len, 0);
        S.inter.assign(len, 0);
        S.queue.clear();
        S.bySize.resize(len);
        std::iota(S.bySize.begin(), S.bySize.end(), ws);
        std::sort(S.bySize.begin(), S.bySize.end(), [&](uint32_t x, uint32_t y) {
            uint32_t dx = g.degree(x), dy = g.degree(y);
            return dx != dy ? dx > dy : x > y;
        });
        uint32_t cursor = 0;
        auto nextSingleton = [&](uint32_t& col) -> bool {
            while (cursor < len && S.placed[S.bySize[cursor] - ws]) ++cursor;
            if (cursor == len) return false;
            col = S.bySize[cursor];
            return true;
        };
        uint32_t unionSize = 0;
        // Adds every row of column `col` not yet in the run's union, and for each such
        // row counts it into the intersection of every unplaced window column holding it.
        auto absorb = [&](uint32_t col) {
            for (uint32_t p = g.colPtrs[col]; p < g.colPtrs[col + 1]; ++p) {
                uint32_t r = g.rows[p];
                if (S.rowMark[r] == S.epoch) continue;
                S.rowMark[r] = S.epoch;
                ++unionSize;
                if (S.rowWin[r] != w + 1) {
                    const uint32_t* b = t->rows.data() + t->colPtrs[r];
                    const uint32_t* e = t->rows.data() + t->colPtrs[r + 1];
                    const uint32_t* lo = std::lower_bound(b, e, ws);
                    const uint32_t* hi = std::lower_bound(lo, e, we);
                    S.rowStart[r] = uint32_t(lo - t->rows.data());
                    S.rowLen[r] = 0;
                    if (uint32_t(hi - lo) >= kLongRun) {
                        S.rowStart[r] = uint32_t(S.pool.size());
                        S.rowLen[r] = kPooled | uint32_t(hi - lo);
                        S.pool.insert(S.pool.end(), lo, hi);
                    }
                    S.rowWin[r] = uint32_t(w) + 1;
                }
                if (S.rowLen[r] & kPooled) {
                    uint32_t* run = S.pool.data() + S.rowStart[r];
                    const uint32_t runLen = S.rowLen[r] & ~kPooled;
                    uint32_t seen = 0;  // placed columns met
                    for (uint32_t i = 0; i < runLen; ++i) {
                        const uint32_t c = run[i];
                        if (S.placed[c - ws]) {
                            ++seen;
                            continue;
                        }
                        if (S.inter[c - ws]++ == 0) S.queue.push_back(c);
                    }
                    if (2 * seen > runLen) {
                        uint32_t j = 0;
                        for (uint32_t i = 0; i < runLen; ++i)
                            if (!S.placed[run[i] - ws]) run[j++] = run[i];
                        S.rowLen[r] = kPooled | j;
                    }
                    continue;
                }
                // placed columns stay placed for the rest of the window: a leading run of
                // them is skipped for good by advancing the row's start
                uint32_t q = S.rowStart[r];
                const uint32_t qe = t->colPtrs[r + 1];
                while (q < qe && t->rows[q] < we && S.placed[t->rows[q] - ws]) ++q;
                S.rowStart[r] = q;
                for (; q < qe; ++q) {
                    uint32_t c = t->rows[q];
                    if (c >= we) break;
                    if (S.placed[c - ws]) continue;
                    if (S.inter[c - ws]++ == 0) S.queue.push_back(c);
                }
            }
        };
        const uint32_t noSets = (len + sliceSize - 1) / sliceSize;
        for (uint32_t set = 0; set < noSets; ++set) {
            for (uint32_t c : S.queue) S.inter[c - ws] = 0;
            S.queue.clear();
            ++S.epoch;
            unionSize = 0;
            const uint32_t ss = ws + set * sliceSize, se = std::min(ss + sliceSize, we);
            uint32_t col;
            nextSingleton(col);
            S.placed[col - ws] = 1;
            absorb(col);
            inv[col] = ss;
            for (uint32_t nc = ss + 1; nc < se; ++nc) {
                uint32_t best;
                if (S.queue.empty()) {
                    if (!nextSingleton(best)) break;
                } else {
                    uint32_t bi = 0;
                    double bv = -1.0;
                    for (uint32_t qi = 0; qi < S.queue.size(); ++qi) {
                        uint32_t c = S.queue[qi];
                        uint32_t inter = S.inter[c - ws];
                        uint32_t uni = unionSize + g.degree(c) - inter;
                        double v = uni == 0 ? 1.0 : double(inter) / double(uni);
                        if (v > bv) {
                            bv = v;
                            bi = qi;
                        }
                    }
                    best = S.queue[bi];
                    S.inter[best - ws] = 0;
                    S.queue[bi] = S.queue.back();
                    S.queue.pop_back();
                }
                S.placed[best - ws] = 1;
                absorb(best);
                inv[best] = nc;
            }
        }
    });
    return inv;
}

// ---------------------------------------------------------------------- Gorder
namespace {

// Max-priority over integer keys with unit updates: one doubly linked list per key
// and a pointer to the largest non-empty key.  Increments move the vertex at once;
// decrements are only counted and applied when the vertex reaches the top (a vertex
// whose stored key is the largest and has no pending decrements holds the true
// maximum), so the window-eviction half of the updates costs one counter write.  A
// vertex's key, pending count and list links share one 16-byte node (the updates are
// random accesses over all n vertices: one cache line each instead of one per field),
// and prefetch() lets a caller that knows its next targets start their loads early.
class BucketQueue {
public:
    explicit BucketQueue(uint32_t n) : nd_(n), head_(16, -1) {
        for (uint32_t v = 0; v < n; ++v) {
            nd_[v].key = 0;
            nd_[v].pend = 0;
            nd_[v].prev = int32_t(v) - 1;
            nd_[v].next = (v + 1 < n) ? int32_t(v) + 1 : -1;
        }
        if (n) head_[0] = 0;
        top_ = 0;
        size_ = n;
    }
    bool contains(uint32_t v) const { return !(nd_[v].pend & kOut); }
    uint64_t size() const { return size_; }
    void prefetch(uint32_t v) const { __builtin_prefetch(&nd_[v], 1, 1); }
    void remove(uint32_t v) {
        unlink(v);
        nd_[v].pend |= kOut;
        --size_;
    }
    void inc(uint32_t v) {
        Node& x = nd_[v];
        if (x.pend & kOut) return;
        unlink(v);
        ++x.key;
        link(v);
        if (x.key > top_) top_ = x.key;
    }
    void dec(uint32_t v) {
        Node& x = nd_[v];
        if (!(x.pend & kOut)) ++x.pend;
    }
    // Largest effective key; among equal keys the most recently linked vertex.
    int64_t pop_max() {
        for (;;) {
            while (top_ > 0 && head_[top_] < 0) --top_;
            const int32_t v = head_[top_];
            if (v < 0) return -1;
            Node& x = nd_[v];
            if (x.pend) {
                unlink(uint32_t(v));
                x.key -= x.pend;
                x.pend = 0;
                link(uint32_t(v));
                continue;
            }
            remove(uint32_t(v));
            return v;
        }
    }

private:
    static constexpr uint32_t kOut = 0x80000000u;  // pend's top bit: not in the queue
    struct Node {
        uint32_t key, pend;
        int32_t prev, next;
    };
    void unlink(uint32_t v) {
        const Node& x = nd_[v];
        if (x.prev >= 0) nd_[x.prev].next = x.next;
        else head_[x.key] = x.next;
        if (x.next >= 0) nd_[x.next].prev = x.prev;
    }
    void link(uint32_t v) {
        Node& x = nd_[v];
        const uint32_t k = x.key;
        if (k >= head_.size()) head_.resize(std::max<size_t>(size_t(k) + 1, head_.size() * 2), -1);
        x.prev = -1;
        x.next = head_[k];
        if (head_[k] >= 0) nd_[head_[k]].prev = int32_t(v);
        head_[k] = int32_t(v);
    }
    std::vector<Node> nd_;
    std::vector<int32_t> head_;
    uint32_t top_;
    uint64_t size_;
};

// The same queue cut into T partitions (blocks of 16 vertices), each with its own
// bucket lists, so that T threads apply the updates of one touch to disjoint nodes at
// once.  Each link carries a stamp from one global sequence (an update's position in the
// sequential update stream, or the moment of a relink at the top), so the most recently
// linked vertex among equal keys is the head of largest stamp over the partitions' top
// buckets: the choice the single queue makes, hence the same order.
class SplitQueue {
public:
    SplitQueue(uint32_t n, unsigned parts) : nd_(n), part_(parts), T_(parts) {
        for (auto& p : part_) {
            p.head.assign(16, -1);
            p.top = 0;
        }
        // bucket 0 holds every vertex, the lowest ID most recent (as the single queue
        // links them): stamp n - v
        std::vector<int32_t> last(T_, -1);
        for (uint32_t v = 0; v < n; ++v) {
            Node& x = nd_[v];
            x.key = 0;
            x.pend = 0;
            x.stamp = uint64_t(n) - v;
            const unsigned p = part_of(v);
            x.prev = last[p];
            x.next = -1;
            if (last[p] >= 0) nd_[last[p]].next = int32_t(v);
            else part_[p].head[0] = int32_t(v);
            last[p] = int32_t(v);
        }
        size_ = n;
    }
    // partitions own blocks of 16 consecutive vertices, so that a cache line of nodes is
    // written by one thread only; T is a power of two
    unsigned part_of(uint32_t v) const { return (v >> 4) & (T_ - 1); }
    bool contains(uint32_t v) const { return !(nd_[v].pend & kOut); }
    uint64_t size() const { return size_; }
    void prefetch(uint32_t v) const { __builtin_prefetch(&nd_[v], 1, 1); }
    void remove(uint32_t v) {
        unlink(v);
        nd_[v].pend |= kOut;
        --size_;
    }
    // only the thread of v's partition calls these during a touch
    void inc(uint32_t v, uint64_t stamp) {
        Node& x = nd_[v];
        if (x.pend & kOut) return;
        unlink(v);
        ++x.key;
        link(v, stamp);
        Part& p = part_[part_of(v)];
        if (x.key > p.top) p.top = x.key;
    }
    void dec(uint32_t v) {
        Node& x = nd_[v];
        if (!(x.pend & kOut)) ++x.pend;
    }
    // Largest effective key; among equal keys the most recently linked vertex.  A relink
    // takes its stamp from `stamp`, which it advances.
    int64_t pop_max(uint64_t& stamp) {
        for (;;) {
            int32_t v = -1;
            uint32_t key = 0;
            for (auto& p : part_) {
                while (p.top > 0 && p.head[p.top] < 0) --p.top;
                const int32_t h = p.head[p.top];
                if (h < 0) continue;
                if (v < 0 || p.top > key || (p.top == key && nd_[h].stamp > nd_[v].stamp)) {
                    v = h;
                    key = p.top;
                }
            }
            if (v < 0) return -1;
            Node& x = nd_[v];
            if (x.pend) {
                unlink(uint32_t(v));
                x.key -= x.pend;
                x.pend = 0;
                link(uint32_t(v), stamp++);
                continue;
            }
            remove(uint32_t(v));
            return v;
        }
    }

private:
    static constexpr uint32_t kOut = 0x80000000u;
    struct Node {
        uint32_t key, pend;
        int32_t prev, next;
        uint64_t stamp;
    };
    struct Part {
        std::vector<int32_t> head;
        uint32_t top;
        char pad[64];  // partitions written by different threads keep to their own lines
    };
    void unlink(uint32_t v) {
        const Node& x = nd_[v];
        if (x.prev >= 0) nd_[x.prev].next = x.next;
        else part_[part_of(v)].head[x.key] = x.next;
        if (x.next >= 0) nd_[x.next].prev = x.prev;
    }
    void link(uint32_t v, uint64_t stamp) {
        Node& x = nd_[v];
        std::vector<int32_t>& head = part_[part_of(v)].head;
        const uint32_t k = x.key;
        if (k >= head.size()) head.resize(std::max<size_t>(size_t(k) + 1, head.size() * 2), -1);
        x.prev = -1;
        x.next = head[k];
        x.stamp = stamp;
        if (head[k] >= 0) nd_[head[k]].prev = int32_t(v);
        head[k] = int32_t(v);
    }
    std::vector<Node> nd_;
    std::vector<Part> part_;
    unsigned T_;
    uint64_t size_;
};

}  // namespace

// Gorder greedy (Wei, Yu, Lu, Lin, "Speedup Graph Processing by Graph Ordering",
// SIGMOD 2016; vendored by the reference at Gorder/Graph.cpp:388-617).  Places the
// vertex maximising the Gorder score S(v, w) = |in(v) & in(w)| + [v->w] + [w->v]
// summed over the last `window` placed vertices w.  As in the reference, in-neighbours
// whose out-degree exceeds sqrt(n) contribute no sibling score, the first vertex is
// the one of largest in-degree, and vertices without edges go last.  Every extraction
// returns a vertex of exactly maximal score (the reference's UnitHeap halves pending
// decrements when it re-files the top).  Unlike the reference (CSC.cuh:936-943) the
// order is returned in the caller's vertex IDs.
std::vector<uint32_t> order_gorder(const CSC& g, uint32_t window, unsigned threads) {
    const uint32_t n = g.n;
    std::vector<uint32_t> inv(n);
    if (n == 0) return inv;
    if (window == 0) window = 1;
    // out(j) = column j's rows (edges j -> i); in(i) = transpose.
    CSC tr;
    const CSC* in = &g;
    if (!csc_is_symmetric(g)) {
        tr = csc_transpose(g);
        in = &tr;
    }
    auto outDeg = [&](uint32_t v) { return g.colPtrs[v + 1] - g.colPtrs[v]; };
    auto inDeg = [&](uint32_t v) { return in->colPtrs[v + 1] - in->colPtrs[v]; };
    const uint64_t huge = uint64_t(std::sqrt(double(n)));
    // The score changes of a touch of v (v entering (+1) or leaving (-1) the window), as a
    // stream of targets: v's out-neighbours, then for each in-neighbour u (not a hub) u
    // itself and u's out-neighbours.  walk() calls f(x, i) for the i-th target x of the
    // stream (i counts every target, so stream positions are the same whoever walks it)
    // and returns the stream's length.  The targets of a run are known ahead: their
    // nodes are prefetched kAhead updates before they are touched.
    constexpr uint32_t kAhead = 8;
    // walk_part: the part of the stream from in-neighbour entries [pA, pB) of v, after
    // v's own out-list when withOut (the stream is the concatenation of such parts).
    auto walk_part = [&](uint32_t v, bool withOut, uint32_t pA, uint32_t i1, auto&& f, auto&& pre) -> uint64_t {
        uint64_t i = 0;
        auto run = [&](uint32_t p0, uint32_t p1) {
            for (uint32_t p = p0; p < p1; ++p) {
                if (p + kAhead < p1) pre(g.rows[p + kAhead]);
                f(g.rows[p], i++);
            }
        };
        if (withOut && outDeg(v) <= huge) run(g.colPtrs[v], g.colPtrs[v + 1]);
        for (uint32_t p = pA; p < i1; ++p) {
            if (p + kAhead < i1) {
                const uint32_t ua = in->rows[p + kAhead];
                pre(ua);
                __builtin_prefetch(&g.colPtrs[ua], 0, 1);
            }
            uint32_t u = in->rows[p];
            if (outDeg(u) > huge) continue;
            f(u, i++);
            if (outDeg(u) > 1) run(g.colPtrs[u], g.colPtrs[u + 1]);
        }
        return i;
    };
    auto walk = [&](uint32_t v, auto&& f, auto&& pre) -> uint64_t {
        return walk_part(v, true, in->colPtrs[v], in->colPtrs[v + 1], f, pre);
    };
    std::vector<uint32_t> order, zero;
    order.reserve(n);
    auto first_vertex = [&](auto& pq) {
        for (uint32_t v = 0; v < n; ++v)
            if (outDeg(v) + inDeg(v) == 0) {
                zero.push_back(v);
                pq.remove(v);
            }
        uint32_t first = 0;
        int64_t best = -1;
        for (uint32_t v = 0; v < n; ++v)
            if (pq.contains(v) && int64_t(inDeg(v)) > best) {
                best = inDeg(v);
                first = v;
            }
        return first;
    };
    threads = std::max(1u, threads);
    if (threads == 1 || n < (1u << 16)) {
        BucketQueue pq(n);
        auto touch = [&](uint32_t v, bool add) {
            walk(
                v, [&](uint32_t x, uint64_t) { add ? pq.inc(x) : pq.dec(x); },
                [&](uint32_t x) { pq.prefetch(x); });
        };
        const uint32_t first = first_vertex(pq);
        if (pq.size() > 0) {
            pq.remove(first);
            order.push_back(first);
            touch(first, true);
            // v is chosen against the last `window` placed vertices; then the vertex that
            // falls out of the window is retired and v's contributions are added.
            while (pq.size() > 0) {
                int64_t v = pq.pop_max();
                order.push_back(uint32_t(v));
                if (order.size() > window) touch(order[order.size() - 1 - window], false);
                touch(uint32_t(v), true);
            }
        }
    } else {
        // The same greedy with the queue split over the threads: a touch with many
        // in-neighbours is walked in slices and its updates dealt to the partitions, each
        // thread applying its own in stream order; a small touch is applied by the calling
        // thread alone.  The random node updates are the cost (~3.5e10 at R-MAT scale
        // 22), and they spread over the threads' memory accesses.
        // Threads: a power of two, at most kMaxSplit (the workers wait for each touch by
        // spinning, which pays only while every one of them has a core).  A touch's stream
        // positions fit 32 bits: a stream is at most the graph's edges plus 2n.
        constexpr unsigned kMaxSplit = 8;
        threads = std::min(threads, kMaxSplit);
        while (threads & (threads - 1)) threads &= threads - 1;
        SplitQueue pq(n, threads);
        uint64_t stamp = uint64_t(n) + 1;
        constexpr uint32_t kSplitIn = 64;  // in-neighbours from which a touch is shared
        // a shared touch runs in two phases, each thread taking one share of each:
        //   walk   thread w walks a contiguous slice of v's in-neighbours (thread 0 first
        //          v's own out-list) and deals each update, with its position in the
        //          slice, to buffer [w][partition of the target];
        //   apply  thread t applies buffers [0][t], [1][t], ... in that order, the
        //          positions offset by the lengths of the slices before, i.e. every update
        //          of its partition in stream order with its stream position as stamp.
        std::vector<std::vector<std::vector<std::pair<uint32_t, uint32_t>>>> buf(
            threads, std::vector<std::vector<std::pair<uint32_t, uint32_t>>>(threads));
        std::vector<uint64_t> sliceLen(threads);
        std::atomic<uint64_t> epoch{0};
        std::atomic<bool> stop{false};
        uint32_t jobV = 0;
        bool jobAdd = false;
        uint64_t jobBase = 0;
        auto spin_wait = [](auto&& ready) {
            for (unsigned spins = 0; !ready();)
                if (++spins >= 4096) {
                    spins = 0;
                    std::this_thread::yield();
                }
        };
        // all threads (the caller included) meet here; `met` counts the thread's own
        // meetings so far (every thread has as many), so the counter never needs a reset
        std::atomic<uint64_t> arrivedAll{0};
        auto meet = [&](uint64_t& met) {
            ++met;
            arrivedAll.fetch_add(1, std::memory_order_acq_rel);
            const uint64_t target = uint64_t(threads) * met;
            spin_wait([&] { return arrivedAll.load(std::memory_order_acquire) >= target; });
        };
        auto walk_share = [&](unsigned w) {
            const uint32_t a0 = in->colPtrs[jobV], a1 = in->colPtrs[jobV + 1];
            const uint32_t pA = a0 + uint32_t(uint64_t(a1 - a0) * w / threads);
            const uint32_t pB = a0 + uint32_t(uint64_t(a1 - a0) * (w + 1) / threads);
            auto& mine = buf[w];
            for (auto& b : mine) b.clear();
            sliceLen[w] = walk_part(
                jobV, w == 0, pA, pB, [&](uint32_t x, uint64_t i) { mine[pq.part_of(x)].emplace_back(x, uint32_t(i)); },
                [](uint32_t) {});
        };
        auto apply_share = [&](unsigned t) {
            uint64_t off = jobBase;
            for (unsigned w = 0; w < threads; ++w) {
                const auto& b = buf[w][t];
                const size_t m = b.size();
                for (size_t k = 0; k < m; ++k) {
                    if (k + 8 < m) pq.prefetch(b[k + 8].first);
                    jobAdd ? pq.inc(b[k].first, off + b[k].second) : pq.dec(b[k].first);
                }
                off += sliceLen[w];
            }
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < threads; ++t)
            pool.emplace_back([&, t] {
                uint64_t seen = 0, met = 0;
                for (;;) {
                    uint64_t e = seen;
                    spin_wait([&] {
                        e = epoch.load(std::memory_order_acquire);
                        return e != seen || stop.load(std::memory_order_relaxed);
                    });
                    if (e == seen) return;  // stopped
                    seen = e;
                    walk_share(t);
                    meet(met);
                    apply_share(t);
                    meet(met);
                }
            });
        uint64_t callerMet = 0;
        auto touch = [&](uint32_t v, bool add) {
            if (inDeg(v) < kSplitIn) {
                stamp += walk(
                    v,
                    [&](uint32_t x, uint64_t i) { add ? pq.inc(x, stamp + i) : pq.dec(x); },
                    [&](uint32_t x) { pq.prefetch(x); });
                return;
            }
            jobV = v;
            jobAdd = add;
            jobBase = stamp;
            epoch.fetch_add(1, std::memory_order_release);
            walk_share(0);
            meet(callerMet);
            apply_share(0);
            meet(callerMet);
            for (unsigned w = 0; w < threads; ++w) stamp += sliceLen[w];
        };
        try {
            const uint32_t first = first_vertex(pq);
            if (pq.size() > 0) {
                pq.remove(first);
                order.push_back(first);
                touch(first, true);
                while (pq.size() > 0) {
                    int64_t v = pq.pop_max(stamp);
                    order.push_back(uint32_t(v));
                    if (order.size() > window) touch(order[order.size() - 1 - window], false);
                    touch(uint32_t(v), true);
                }
            }
        } catch (...) {
            stop.store(true);
            for (auto& th : pool) th.join();
            throw;
        }
        stop.store(true);
        for (auto& th : pool) th.join();
    }
    order.insert(order.end(), zero.begin(), zero.end());
    for (uint32_t k = 0; k < n; ++k) inv[order[k]] = k;
    return inv;
}

std::vector<uint32_t> chain_permutations(const std::vector<uint32_t>& first,
                                         const std::vector<uint32_t>& second) {
    std::vector<uint32_t> out(first.size());
    for (size_t i = 0; i < first.size(); ++i) out[i] = second[first[i]];
    return out;
}

std::vector<uint32_t> reorder(CSC& g, const blest_reorder_opts& o) {
    const uint32_t slice = o.slice_size ? o.slice_size : 8;
    const unsigned threads = hw_threads(o.threads);
    blest_reorder_kind kind = o.kind;
    if (kind == BLEST_REORDER_AUTO) {  // CSC::reorder (CSC.cuh:358-386)
        if (g.isSocial)
            kind = o.jaccard ? (o.gorder ? BLEST_REORDER_GORDER_JACCARD : BLEST_REORDER_JACCARD)
                             : BLEST_REORDER_NATURAL;
        else
            kind = BLEST_REORDER_RCM;
    }
    std::vector<uint32_t> inv;
    switch (kind) {
        case BLEST_REORDER_NATURAL: inv = order_natural(g); break;
        case BLEST_REORDER_RANDOM: inv = order_random(g, o.seed); break;
        case BLEST_REORDER_DEGREE: inv = order_degree(g, true); break;
        case BLEST_REORDER_RCM: inv = order_rcm(g, o.seed); break;
        case BLEST_REORDER_JACCARD: inv = order_jaccard_window(g, slice, o.window, threads); break;
        case BLEST_REORDER_GORDER: inv = order_gorder(g, o.window, threads); break;
        case BLEST_REORDER_GORDER_JACCARD: {  // CSC::gorderWithJackard (CSC.cuh:932-948)
            std::vector<uint32_t> first = order_gorder(g, o.window, threads);
            csc_apply_permutation(g, first.data(), threads);
            std::vector<uint32_t> second = order_jaccard_window(g, slice, o.window, threads);
            csc_apply_permutation(g, second.data(), threads);
            return chain_permutations(first, second);
        }
        default: throw Error(BLEST_ERR_INVALID_ARGUMENT, "unknown reorder kind");
    }
    csc_apply_permutation(g, inv.data(), threads);
    return inv;
}

}  // namespace blest
