// CSC construction, I/O, generators and graph utilities (host).
#include "csc.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <numeric>

namespace blest {

namespace {

inline void atomic_inc(uint32_t* p) { __atomic_fetch_add(p, 1u, __ATOMIC_RELAXED); }
inline uint64_t atomic_take(uint64_t* p) { return __atomic_fetch_add(p, uint64_t(1), __ATOMIC_RELAXED); }

// Sort + unique every column of a (ptr, rows) pattern held in `tmp` with u64 offsets,
// drop self loops, and compact into g.
void compact_columns(uint32_t n, const std::vector<uint64_t>& off, uint32_t* tmp, CSC& g,
                     unsigned threads) {
    std::vector<uint32_t> newDeg(n, 0);
    parallel_dynamic(n, threads, 4096, [&](uint64_t j, unsigned) {
        uint32_t* b = tmp + off[j];
        uint32_t* e = tmp + off[j + 1];
        std::sort(b, e);
        uint32_t* w = b;
        for (uint32_t* r = b; r != e; ++r) {
            if (*r == uint32_t(j)) continue;              // self loop (CSC.cuh:125)
            if (w != b && *(w - 1) == *r) continue;       // duplicate
            *w++ = *r;
        }
        newDeg[j] = uint32_t(w - b);
    });
    uint64_t total = 0;
    for (uint32_t j = 0; j < n; ++j) total += newDeg[j];
    if (total > 0xFFFFFFFFull)
        throw Error(BLEST_ERR_INVALID_ARGUMENT, "graph has >= 2^32 directed edges; CSC offsets are u32");
    g.n = n;
    g.colPtrs.assign(size_t(n) + 1, 0);
    for (uint32_t j = 0; j < n; ++j) g.colPtrs[j + 1] = g.colPtrs[j] + newDeg[j];
    g.rows.resize(total);
    parallel_dynamic(n, threads, 4096, [&](uint64_t j, unsigned) {
        std::copy(tmp + off[j], tmp + off[j] + newDeg[j], g.rows.data() + g.colPtrs[j]);
    });
}

}  // namespace

void csc_normalise(CSC& g, unsigned threads) {
    threads = hw_threads(threads);
    std::vector<uint64_t> off(size_t(g.n) + 1);
    for (size_t i = 0; i <= g.n; ++i) off[i] = g.colPtrs[i];
    std::vector<uint32_t> tmp(std::move(g.rows));
    compact_columns(g.n, off, tmp.data(), g, threads);
    g.isSocial = csc_is_social(g);
}

CSC csc_from_arrays(uint32_t n, const uint32_t* colPtrs, const uint32_t* rows) {
    if (!colPtrs) throw Error(BLEST_ERR_INVALID_ARGUMENT, "col_ptrs is null");
    if (colPtrs[0] != 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "col_ptrs[0] != 0");
    for (uint32_t j = 0; j < n; ++j)
        if (colPtrs[j + 1] < colPtrs[j]) throw Error(BLEST_ERR_INVALID_ARGUMENT, "col_ptrs not monotone");
    uint64_t nnz = colPtrs[n];
    if (nnz && !rows) throw Error(BLEST_ERR_INVALID_ARGUMENT, "rows is null");
    for (uint64_t p = 0; p < nnz; ++p)
        if (rows[p] >= n) throw Error(BLEST_ERR_INVALID_ARGUMENT, "row index out of range");
    CSC g;
    g.n = n;
    g.colPtrs.assign(colPtrs, colPtrs + n + 1);
    g.rows.assign(rows, rows + nnz);
    csc_normalise(g);
    return g;
}

// ------------------------------------------------------------------- MTX reader
CSC csc_from_mtx(const std::string& path, bool undirected, bool binary) {
    (void)binary;  // the value column, if any, is skipped to end of line
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(BLEST_ERR_IO, "Failed to open file from which to construct CSC: " + path);
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::string buf(size_t(std::max(0L, sz)), '\0');
    size_t got = sz > 0 ? std::fread(&buf[0], 1, size_t(sz), f) : 0;
    std::fclose(f);
    buf.resize(got);
    const char* p = buf.data();
    const char* end = p + buf.size();
    auto skip_line = [&] {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    };
    auto skip_ws = [&] {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    };
    auto read_u64 = [&](uint64_t& v) -> bool {
        skip_ws();
        if (p >= end || *p < '0' || *p > '9') return false;
        v = 0;
        while (p < end && *p >= '0' && *p <= '9') v = v * 10 + uint64_t(*p++ - '0');
        return true;
    };
    while (p < end && (*p == '%' || *p == '\n')) skip_line();
    uint64_t nr, nc, lines;
    if (!read_u64(nr) || !read_u64(nc) || !read_u64(lines))
        throw Error(BLEST_ERR_IO, "malformed MatrixMarket size line in " + path);
    skip_line();
    if (nc > 0xFFFFFFFEull) throw Error(BLEST_ERR_INVALID_ARGUMENT, "too many vertices");
    uint32_t n = uint32_t(nc);  // CSC.cuh:114 reads rows then cols into m_N
    std::vector<std::pair<uint32_t, uint32_t>> edges;  // (column j, row i): edge j -> i
    edges.reserve(size_t(lines) * (undirected ? 2 : 1));
    for (uint64_t it = 0; it < lines; ++it) {
        while (p < end && (*p == '%' || *p == '\n')) skip_line();
        uint64_t j, i;
        if (!read_u64(j) || !read_u64(i)) throw Error(BLEST_ERR_IO, "truncated MatrixMarket file " + path);
        skip_line();
        if (i == j) continue;
        if (i == 0 || j == 0 || i > n || j > n) throw Error(BLEST_ERR_IO, "MatrixMarket index out of range");
        edges.emplace_back(uint32_t(j - 1), uint32_t(i - 1));
        if (undirected) edges.emplace_back(uint32_t(i - 1), uint32_t(j - 1));
    }
    std::vector<uint64_t> off(size_t(n) + 1, 0);
    for (auto& e : edges) ++off[e.first + 1];
    for (uint32_t j = 0; j < n; ++j) off[j + 1] += off[j];
    std::vector<uint32_t> tmp(edges.size());
    std::vector<uint64_t> cur(off.begin(), off.end() - 1);
    for (auto& e : edges) tmp[cur[e.first]++] = e.second;
    edges.clear();
    edges.shrink_to_fit();
    CSC g;
    compact_columns(n, off, tmp.data(), g, hw_threads());
    g.isSocial = csc_is_social(g);
    return g;
}

// ----------------------------------------------------------------- binary cache
CSC csc_load_binary(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error(BLEST_ERR_IO, "cannot open CSC cache " + path);
    CSC g;
    uint32_t nnz = 0;
    uint8_t social = 0;
    in.read(reinterpret_cast<char*>(&g.n), 4);
    in.read(reinterpret_cast<char*>(&nnz), 4);
    in.read(reinterpret_cast<char*>(&social), 1);
    if (!in) throw Error(BLEST_ERR_IO, "truncated CSC cache " + path);
    g.isSocial = social != 0;
    g.colPtrs.resize(size_t(g.n) + 1);
    g.rows.resize(nnz);
    in.read(reinterpret_cast<char*>(g.colPtrs.data()), std::streamsize(4) * (g.n + 1));
    in.read(reinterpret_cast<char*>(g.rows.data()), std::streamsize(4) * nnz);
    if (!in) throw Error(BLEST_ERR_IO, "truncated CSC cache " + path);
    if (g.colPtrs[g.n] != nnz) throw Error(BLEST_ERR_IO, "inconsistent CSC cache " + path);
    return g;
}

void csc_save_binary(const CSC& g, const std::string& path) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw Error(BLEST_ERR_IO, "cannot write CSC cache " + path);
    uint32_t nnz = uint32_t(g.nnz());
    uint8_t social = g.isSocial ? 1 : 0;
    out.write(reinterpret_cast<const char*>(&g.n), 4);
    out.write(reinterpret_cast<const char*>(&nnz), 4);
    out.write(reinterpret_cast<const char*>(&social), 1);
    out.write(reinterpret_cast<const char*>(g.colPtrs.data()), std::streamsize(4) * (g.n + 1));
    out.write(reinterpret_cast<const char*>(g.rows.data()), std::streamsize(4) * nnz);
    if (!out) throw Error(BLEST_ERR_IO, "write failed " + path);
}

// ------------------------------------------------------------------ edge streams
CSC csc_from_edge_stream(uint32_t n, uint64_t m, EdgeFn fn, const void* ctx, unsigned threads) {
    threads = hw_threads(threads);
    std::unique_ptr<uint32_t[]> cnt(new uint32_t[n]());
    parallel_blocks(m, threads, [&](uint64_t b, uint64_t e, unsigned) {
        uint32_t u, v;
        for (uint64_t k = b; k < e; ++k) {
            if (!fn(ctx, k, &u, &v) || u == v) continue;
            atomic_inc(&cnt[u]);
            atomic_inc(&cnt[v]);
        }
    });
    std::vector<uint64_t> off(size_t(n) + 1, 0);
    for (uint32_t j = 0; j < n; ++j) off[j + 1] = off[j] + cnt[j];
    cnt.reset();
    std::unique_ptr<uint32_t[]> tmp(new uint32_t[std::max<uint64_t>(1, off[n])]);
    std::vector<uint64_t> cur(off.begin(), off.end() - 1);
    parallel_blocks(m, threads, [&](uint64_t b, uint64_t e, unsigned) {
        uint32_t u, v;
        for (uint64_t k = b; k < e; ++k) {
            if (!fn(ctx, k, &u, &v) || u == v) continue;
            tmp[atomic_take(&cur[u])] = v;
            tmp[atomic_take(&cur[v])] = u;
        }
    });
    CSC g;
    compact_columns(n, off, tmp.get(), g, threads);
    tmp.reset();
    g.isSocial = csc_is_social(g);
    return g;
}

std::vector<uint32_t> seeded_permutation(uint32_t n, uint64_t seed, uint64_t stream) {
    std::vector<uint32_t> p(n);
    std::iota(p.begin(), p.end(), 0u);
    uint64_t key = stream_key(seed, stream);
    for (uint32_t i = n ? n - 1 : 0; i > 0; --i) {
        uint32_t j = uint32_t(crand(key, i) % (uint64_t(i) + 1));
        std::swap(p[i], p[j]);
    }
    return p;
}

namespace {

// Stream ids of the generators' counter-based draws.
enum : uint64_t {
    kStreamRmat = 1,
    kStreamRmatPerm = 2,
    kStreamGridRight = 3,
    kStreamGridDown = 4,
    kStreamGridShortcut = 5,
    kStreamEr = 6,
    kStreamRoots = 7
};

struct RmatCtx {
    uint32_t scale;
    double a, ab, abc;
    uint64_t key;
    const uint32_t* perm;
};

// One sample: for each of the `scale` recursion levels draw a 32-bit uniform (two
// levels per 64-bit draw: high half first) and descend into quadrant
// a=(0,0), b=(0,1), c=(1,0), d=(1,1) of (u, v).
bool rmat_edge(const void* c, uint64_t k, uint32_t* u, uint32_t* v) {
    const RmatCtx& x = *static_cast<const RmatCtx*>(c);
    const uint32_t pairs = (x.scale + 1) / 2;
    uint32_t su = 0, sv = 0;
    uint64_t h = 0;
    for (uint32_t l = 0; l < x.scale; ++l) {
        uint32_t half;
        if ((l & 1) == 0) {
            h = crand(x.key, k * pairs + l / 2);
            half = uint32_t(h >> 32);
        } else {
            half = uint32_t(h);
        }
        double r = double(half) * (1.0 / 4294967296.0);
        uint32_t bu, bv;
        if (r < x.a) { bu = 0; bv = 0; }
        else if (r < x.ab) { bu = 0; bv = 1; }
        else if (r < x.abc) { bu = 1; bv = 0; }
        else { bu = 1; bv = 1; }
        su = (su << 1) | bu;
        sv = (sv << 1) | bv;
    }
    if (x.perm) {
        su = x.perm[su];
        sv = x.perm[sv];
    }
    *u = su;
    *v = sv;
    return true;
}

struct GridCtx {
    uint32_t side;
    uint64_t n, shortcuts;
    uint64_t keep32;  // keep threshold on a 32-bit draw
    uint64_t kr, kd, ks;
};

bool grid_edge(const void* c, uint64_t k, uint32_t* u, uint32_t* v) {
    const GridCtx& x = *static_cast<const GridCtx*>(c);
    if (k < x.n) {  // right edge of vertex k
        uint32_t r = uint32_t(k / x.side), col = uint32_t(k % x.side);
        if (col + 1 >= x.side) return false;
        if ((crand(x.kr, k) >> 32) >= x.keep32) return false;
        *u = uint32_t(k);
        *v = r * x.side + col + 1;
        return true;
    }
    if (k < 2 * x.n) {  // down edge of vertex k - n
        uint64_t id = k - x.n;
        uint32_t r = uint32_t(id / x.side);
        if (r + 1 >= x.side) return false;
        if ((crand(x.kd, id) >> 32) >= x.keep32) return false;
        *u = uint32_t(id);
        *v = uint32_t(id + x.side);
        return true;
    }
    uint64_t s = k - 2 * x.n;  // shortcut s
    *u = uint32_t(crand(x.ks, 2 * s) % x.n);
    *v = uint32_t(crand(x.ks, 2 * s + 1) % x.n);
    return *u != *v;
}

struct ErCtx {
    uint64_t n, key;
};

bool er_edge(const void* c, uint64_t k, uint32_t* u, uint32_t* v) {
    const ErCtx& x = *static_cast<const ErCtx*>(c);
    *u = uint32_t(crand(x.key, 2 * k) % x.n);
    *v = uint32_t(crand(x.key, 2 * k + 1) % x.n);
    return *u != *v;
}

}  // namespace

CSC generate_rmat(uint32_t scale, uint32_t edgeFactor, double a, double b, double c,
                  uint64_t seed, bool permute, unsigned threads) {
    if (scale < 1 || scale > 31) throw Error(BLEST_ERR_INVALID_ARGUMENT, "rmat scale must be in [1, 31]");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        throw Error(BLEST_ERR_INVALID_ARGUMENT, "rmat probabilities must be >= 0 and sum <= 1");
    uint32_t n = 1u << scale;
    std::vector<uint32_t> perm;
    if (permute) perm = seeded_permutation(n, seed, kStreamRmatPerm);
    RmatCtx ctx{scale, a, a + b, a + b + c, stream_key(seed, kStreamRmat), permute ? perm.data() : nullptr};
    return csc_from_edge_stream(n, uint64_t(edgeFactor) << scale, rmat_edge, &ctx, threads);
}

CSC generate_grid(uint32_t side, double keep, double shortcutFrac, uint64_t seed, unsigned threads) {
    if (side < 1 || uint64_t(side) * side > 0xFFFFFFFEull)
        throw Error(BLEST_ERR_INVALID_ARGUMENT, "grid side out of range");
    if (keep < 0 || keep > 1 || shortcutFrac < 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "bad grid probabilities");
    GridCtx ctx;
    ctx.side = side;
    ctx.n = uint64_t(side) * side;
    ctx.shortcuts = uint64_t(std::llround(shortcutFrac * double(ctx.n)));
    ctx.keep32 = keep >= 1.0 ? (1ull << 32) : uint64_t(keep * 4294967296.0);
    ctx.kr = stream_key(seed, kStreamGridRight);
    ctx.kd = stream_key(seed, kStreamGridDown);
    ctx.ks = stream_key(seed, kStreamGridShortcut);
    return csc_from_edge_stream(uint32_t(ctx.n), 2 * ctx.n + ctx.shortcuts, grid_edge, &ctx, threads);
}

CSC generate_er(uint32_t n, double avgDegree, uint64_t seed, unsigned threads) {
    if (n < 1) throw Error(BLEST_ERR_INVALID_ARGUMENT, "er n must be >= 1");
    if (avgDegree < 0) throw Error(BLEST_ERR_INVALID_ARGUMENT, "er degree must be >= 0");
    ErCtx ctx{n, stream_key(seed, kStreamEr)};
    uint64_t m = uint64_t(std::llround(double(n) * avgDegree / 2.0));
    return csc_from_edge_stream(n, m, er_edge, &ctx, threads);
}

// ------------------------------------------------------------------ utilities
CSC csc_transpose(const CSC& g, unsigned threads) {
    threads = hw_threads(threads);
    CSC t;
    t.n = g.n;
    t.isSocial = g.isSocial;
    t.colPtrs.assign(size_t(g.n) + 1, 0);
    for (uint32_t r : g.rows) ++t.colPtrs[r + 1];
    for (uint32_t i = 0; i < g.n; ++i) t.colPtrs[i + 1] += t.colPtrs[i];
    t.rows.resize(g.nnz());
    std::vector<uint32_t> cur(t.colPtrs.begin(), t.colPtrs.end() - 1);
    // Visiting columns in ascending order leaves every transposed column sorted.
    for (uint32_t j = 0; j < g.n; ++j)
        for (uint32_t p = g.colPtrs[j]; p < g.colPtrs[j + 1]; ++p) t.rows[cur[g.rows[p]]++] = j;
    (void)threads;
    return t;
}

CSC csc_symmetrize(const CSC& g, unsigned threads) {
    threads = hw_threads(threads);
    std::vector<uint64_t> off(size_t(g.n) + 1, 0);
    for (uint32_t j = 0; j < g.n; ++j) {
        off[j + 1] += g.degree(j);
        for (uint32_t p = g.colPtrs[j]; p < g.colPtrs[j + 1]; ++p) off[g.rows[p] + 1] += 1;
    }
    for (uint32_t j = 0; j < g.n; ++j) off[j + 1] += off[j];
    std::vector<uint32_t> tmp(off[g.n]);
    std::vector<uint64_t> cur(off.begin(), off.end() - 1);
    for (uint32_t j = 0; j < g.n; ++j)
        for (uint32_t p = g.colPtrs[j]; p < g.colPtrs[j + 1]; ++p) {
            uint32_t i = g.rows[p];
            tmp[cur[j]++] = i;
            tmp[cur[i]++] = j;
        }
    CSC s;
    compact_columns(g.n, off, tmp.data(), s, threads);
    s.isSocial = csc_is_social(s);
    return s;
}

bool csc_is_symmetric(const CSC& g, unsigned threads) {
    threads = hw_threads(threads);
    std::atomic<bool> ok{true};
    parallel_blocks(g.n, threads, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t j = b; j < e && ok.load(std::memory_order_relaxed); ++j)
            for (uint32_t p = g.colPtrs[j]; p < g.colPtrs[j + 1]; ++p) {
                uint32_t i = g.rows[p];
                if (i == j) continue;
                const uint32_t* cb = g.rows.data() + g.colPtrs[i];
                const uint32_t* ce = g.rows.data() + g.colPtrs[i + 1];
                if (!std::binary_search(cb, ce, uint32_t(j))) {
                    ok.store(false);
                    return;
                }
            }
    });
    return ok.load();
}

bool social_degree_distribution(const std::vector<uint32_t>& deg) {
    // Restates Common.cuh:100-250: heavy tail by degree shares OR a power-law tail by
    // least squares on log(count) vs log(degree) for degrees >= 5.
    const uint64_t n = deg.size();
    if (n == 0) return false;
    uint32_t maxDeg = 0;
    uint64_t sumDeg = 0;
    for (uint32_t d : deg) {
        sumDeg += d;
        maxDeg = std::max(maxDeg, d);
    }
    if (maxDeg == 0 || sumDeg == 0) return false;
    std::vector<uint32_t> sorted(deg);
    std::sort(sorted.begin(), sorted.end());
    auto shareTop = [&](double frac) {
        uint64_t k = uint64_t(std::ceil(frac * double(n)));
        k = std::max<uint64_t>(1, std::min<uint64_t>(k, n));
        uint64_t top = 0;
        for (uint64_t i = 0; i < k; ++i) top += sorted[n - 1 - i];
        return double(top) / double(sumDeg);
    };
    const bool byShares = shareTop(0.01) >= 0.05 && shareTop(0.10) >= 0.40;
    std::vector<uint64_t> freq(size_t(maxDeg) + 1, 0);
    for (uint32_t d : deg) ++freq[d];
    std::vector<double> xs, ys;
    for (uint32_t k = 5; k <= maxDeg; ++k)
        if (freq[k]) {
            xs.push_back(std::log(double(k)));
            ys.push_back(std::log(double(freq[k])));
        }
    bool powerLaw = false;
    if (xs.size() >= 3) {
        double sx = 0, sy = 0, sxy = 0, sxx = 0;
        for (size_t i = 0; i < xs.size(); ++i) {
            sx += xs[i];
            sy += ys[i];
            sxy += xs[i] * ys[i];
            sxx += xs[i] * xs[i];
        }
        const double m = double(xs.size());
        const double den = m * sxx - sx * sx;
        if (den > 0) {
            const double slope = (m * sxy - sx * sy) / den;
            const double icpt = (sy - slope * sx) / m;
            const double ym = sy / m;
            double tot = 0, res = 0;
            for (size_t i = 0; i < xs.size(); ++i) {
                const double dt = ys[i] - ym, dr = ys[i] - (icpt + slope * xs[i]);
                tot += dt * dt;
                res += dr * dr;
            }
            if (tot > 0) {
                const double r2 = 1.0 - res / tot, gamma = -slope;
                powerLaw = r2 >= 0.70 && gamma >= 1.0 && gamma <= 5.0;
            }
        }
    }
    return byShares || powerLaw;
}

bool csc_is_social(const CSC& g) {
    if (g.n == 0) return false;
    std::vector<uint32_t> outDeg(g.n), inDeg(g.n, 0);
    for (uint32_t j = 0; j < g.n; ++j) outDeg[j] = g.degree(j);
    for (uint32_t r : g.rows) ++inDeg[r];
    const double avg = double(g.nnz()) / double(g.n);  // CSC::averageDegree (CSC.cuh:324-333)
    return social_degree_distribution(inDeg) || social_degree_distribution(outDeg) || avg > 18.0;
}

bool permutation_is_bijection(const uint32_t* inv, uint32_t n) {
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t j = 0; j < n; ++j) {
        if (inv[j] >= n || seen[inv[j]]) return false;
        seen[inv[j]] = 1;
    }
    return true;
}

void csc_apply_permutation(CSC& g, const uint32_t* inv, unsigned threads) {
    threads = hw_threads(threads);
    if (!inv || !permutation_is_bijection(inv, g.n)) throw Error(BLEST_ERR_PERMUTATION, "Permutation check has failed.");
    std::vector<uint32_t> ptr(size_t(g.n) + 1, 0);
    for (uint32_t old = 0; old < g.n; ++old) ptr[inv[old] + 1] = g.degree(old);
    for (uint32_t j = 0; j < g.n; ++j) ptr[j + 1] += ptr[j];
    std::vector<uint32_t> rows(g.nnz());
    parallel_dynamic(g.n, threads, 2048, [&](uint64_t old, unsigned) {
        uint32_t nc = inv[old];
        uint32_t* dst = rows.data() + ptr[nc];
        for (uint32_t p = g.colPtrs[old]; p < g.colPtrs[old + 1]; ++p) *dst++ = inv[g.rows[p]];
        std::sort(rows.data() + ptr[nc], rows.data() + ptr[nc + 1]);
    });
    g.colPtrs.swap(ptr);
    g.rows.swap(rows);
}

CSCMetrics csc_metrics(const CSC& g) {
    CSCMetrics m{0, 0, 0, 0, 0, 0};
    double sb = 0, sp = 0, sd = 0;
    for (uint32_t j = 0; j < g.n; ++j) {
        uint32_t b = g.colPtrs[j], e = g.colPtrs[j + 1];
        sd += e - b;
        m.maxDegree = std::max(m.maxDegree, e - b);
        if (b != e) {
            uint32_t band = uint32_t(std::abs(int64_t(j) - int64_t(g.rows[e - 1])));
            m.maxBandwidth = std::max(m.maxBandwidth, band);
            sb += band;
        }
        if (b + 2 <= e) {
            uint32_t prof = g.rows[e - 1] - g.rows[b];
            m.maxProfile = std::max(m.maxProfile, prof);
            sp += prof;
        }
    }
    if (g.n) {
        m.avgBandwidth = sb / g.n;
        m.avgProfile = sp / g.n;
        m.avgDegree = sd / g.n;
    }
    return m;
}

std::vector<uint32_t> pick_roots(const CSC& g, uint32_t count, uint64_t seed) {
    uint64_t eligible = 0;
    for (uint32_t j = 0; j < g.n; ++j) eligible += g.degree(j) > 0;
    if (count > eligible) throw Error(BLEST_ERR_INVALID_ARGUMENT, "not enough vertices of degree >= 1 for the requested roots");
    std::vector<uint32_t> roots;
    std::vector<uint8_t> taken(g.n, 0);
    uint64_t key = stream_key(seed, kStreamRoots);
    for (uint64_t t = 0; roots.size() < count; ++t) {
        uint32_t v = uint32_t(crand(key, t) % g.n);
        if (g.degree(v) == 0 || taken[v]) continue;
        taken[v] = 1;
        roots.push_back(v);
    }
    return roots;
}

}  // namespace blest
