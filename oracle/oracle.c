/*
 * oracle.c — CPU restatement of the reference traversal path.
 *
 * TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py, as the checker or the timed baseline.  The product path
 * (libblest.so) never links, loads or calls it.
 *
 * Each function names the reference lines it restates.  How each one is pinned
 * (textbook closed forms, scipy, brute force, hand-computed fixtures) is listed in
 * the test files under tests/.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define NOV 0xFFFFFFFFu

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------- textbook BFS */
uint32_t oracle_bfs(uint32_t n, const uint32_t* cp, const uint32_t* rows, uint32_t src, int32_t* levels) {
    for (uint32_t v = 0; v < n; ++v) levels[v] = -1;
    if (src >= n) return 0;
    uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t qs = 0, qe = 0;
    q[qe++] = src;
    levels[src] = 0;
    while (qs < qe) {
        uint32_t u = q[qs++];
        for (uint32_t p = cp[u]; p < cp[u + 1]; ++p) {
            uint32_t v = rows[p];
            if (levels[v] < 0) {
                levels[v] = levels[u] + 1;
                q[qe++] = v;
            }
        }
    }
    free(q);
    return qe;
}

/* ------------------------------------------------------------------ union-find */
static uint32_t uf_find(uint32_t* p, uint32_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

uint32_t oracle_cc(uint32_t n, const uint32_t* cp, const uint32_t* rows, uint32_t* labels) {
    uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) p[v] = v;
    for (uint32_t j = 0; j < n; ++j)
        for (uint32_t k = cp[j]; k < cp[j + 1]; ++k) {
            uint32_t a = uf_find(p, j), b = uf_find(p, rows[k]);
            if (a == b) continue;
            if (a < b) p[b] = a; /* the smaller root survives */
            else p[a] = b;
        }
    uint32_t comps = 0;
    for (uint32_t v = 0; v < n; ++v) {
        labels[v] = uf_find(p, v);
        comps += labels[v] == v;
    }
    free(p);
    return comps;
}

/* ------------------------------------------------------------ BVSS construction */
/* Merge of the rset's columns exactly as BVSS.cuh:205-252: advance every column
 * pointer past rows < i, collect the columns whose next row is i, jump i to the
 * smallest next row.  Writes up to cap (row, mask) pairs; returns the slice count. */
static uint64_t merge_rset(uint32_t n, const uint32_t* cp, const uint32_t* rows, uint32_t s, uint32_t rset,
                           uint32_t* outRow, uint32_t* outMask) {
    uint32_t c0 = rset * s;
    uint32_t c1 = c0 + s < n ? c0 + s : n;
    uint32_t k = c1 - c0;
    uint32_t ptr[32];
    for (uint32_t idx = 0; idx < k; ++idx) ptr[idx] = cp[c0 + idx];
    uint64_t count = 0;
    uint32_t i = 0;
    while (i < n) {
        uint32_t individual = 0, nextRow = n;
        for (uint32_t idx = 0; idx < k; ++idx) {
            uint32_t j = c0 + idx;
            while (ptr[idx] < cp[j + 1] && rows[ptr[idx]] < i) ++ptr[idx];
            uint32_t p = ptr[idx];
            if (p < cp[j + 1]) {
                if (rows[p] == i) {
                    individual |= 1u << idx;
                    ++p;
                }
                ptr[idx] = p;
                if (p < cp[j + 1] && rows[p] < nextRow) nextRow = rows[p];
            }
        }
        if (individual) {
            if (outRow) {
                outRow[count] = i;
                outMask[count] = individual;
            }
            ++count;
        }
        i = nextRow;
    }
    return count;
}

int oracle_bvss_build(uint32_t n, const uint32_t* cp, const uint32_t* rows, uint32_t s, uint32_t nm, int fullPad,
                      uint32_t* out_v, uint64_t* out_s, uint64_t* ssp, uint32_t* v2r, uint32_t* rp,
                      uint32_t* rowIds, uint32_t* masks, double* div) {
    if (s == 0 || s > 32 || 128 % s != 0) return -1; /* BVSS.cuh:103-106 */
    if (nm != 32 / s) return -1;                     /* BVSS.cuh:107-110 */
    const uint32_t R = (uint32_t)(((uint64_t)n + s - 1) / s);
    const uint64_t full = 32ull * nm; /* WARP_SIZE * noMasks (BVSS.cuh:328) */
    uint32_t maxCol = 0;
    for (uint32_t j = 0; j < n; ++j)
        if (cp[j + 1] - cp[j] > maxCol) maxCol = cp[j + 1] - cp[j];
    uint32_t* tr = (uint32_t*)malloc(sizeof(uint32_t) * ((uint64_t)maxCol * s + 1));
    uint32_t* tm = (uint32_t*)malloc(sizeof(uint32_t) * ((uint64_t)maxCol * s + 1));
    uint32_t V = 0;
    uint64_t S = 0;
    double divSum = 0;
    if (rp) rp[0] = 0;
    if (ssp) ssp[0] = 0;
    for (uint32_t r = 0; r < R; ++r) {
        uint64_t cnt = merge_rset(n, cp, rows, s, r, tr, tm);
        uint64_t complete = cnt / full;
        uint64_t leftStart = complete * full;
        uint32_t nv = (uint32_t)(complete + (leftStart < cnt ? 1 : 0));
        for (uint32_t vi = 0; vi < nv; ++vi) {
            /* vset vi of this rset: list of (lane, position, source index or pad) */
            uint64_t base = S;
            uint64_t size;
            if (vi < complete) {
                size = full;
                if (ssp) /* BVSS.cuh:335-346 */
                    for (uint32_t t = 0; t < 32; ++t) {
                        uint32_t cum = 0;
                        for (uint32_t m = 0; m < nm; ++m) {
                            uint64_t cur = vi * full + (uint64_t)m * 32 + t;
                            rowIds[base + (uint64_t)t * nm + m] = tr[cur];
                            cum |= tm[cur] << (s * m);
                        }
                        masks[base / nm + t] = cum;
                    }
            } else if (fullPad) {
                size = full;
                if (ssp) /* BVSS.cuh:356-378 */
                    for (uint32_t t = 0; t < 32; ++t) {
                        uint32_t cum = 0;
                        for (uint32_t m = 0; m < nm; ++m) {
                            uint64_t cur = leftStart + (uint64_t)m * 32 + t;
                            if (cur < cnt) {
                                rowIds[base + (uint64_t)t * nm + m] = tr[cur];
                                cum |= tm[cur] << (s * m);
                            } else {
                                rowIds[base + (uint64_t)t * nm + m] = NOV;
                            }
                        }
                        masks[base / nm + t] = cum;
                    }
            } else { /* BVSS.cuh:379-415 */
                uint64_t stride = (cnt - leftStart) / nm;
                uint64_t rem = (cnt - leftStart) - stride * nm;
                size = stride * nm + (rem ? nm : 0);
                if (ssp) {
                    for (uint64_t t = 0; t < stride; ++t) {
                        uint32_t cum = 0;
                        for (uint32_t m = 0; m < nm; ++m) {
                            uint64_t cur = leftStart + (uint64_t)m * stride + t;
                            rowIds[base + t * nm + m] = tr[cur];
                            cum |= tm[cur] << (s * m);
                        }
                        masks[base / nm + t] = cum;
                    }
                    if (rem) {
                        uint64_t ls = leftStart + stride * nm;
                        uint32_t cum = 0;
                        for (uint32_t m = 0; m < nm; ++m) {
                            uint64_t cur = ls + m;
                            if (cur < cnt) {
                                rowIds[base + stride * nm + m] = tr[cur];
                                cum |= tm[cur] << (s * m);
                            } else {
                                rowIds[base + stride * nm + m] = NOV;
                            }
                        }
                        masks[base / nm + stride] = cum;
                    }
                }
            }
            if (ssp) {
                /* computeUpdateDivergence term (BVSS.cuh:424-462) */
                double vmean = 0;
                for (uint32_t m = 0; m < nm; ++m) {
                    double mean = 0, sd = 0;
                    uint32_t c = 0;
                    for (uint32_t t = 0; t < 32; ++t) {
                        uint64_t sl = (uint64_t)t * nm + m;
                        if (sl < size && rowIds[base + sl] != NOV) {
                            mean += rowIds[base + sl];
                            ++c;
                        }
                    }
                    if (c) mean /= c;
                    for (uint32_t t = 0; t < 32; ++t) {
                        uint64_t sl = (uint64_t)t * nm + m;
                        if (sl < size && rowIds[base + sl] != NOV) {
                            double d = rowIds[base + sl] - mean;
                            sd += d * d;
                        }
                    }
                    if (c) sd /= c;
                    vmean += sqrt(sd);
                }
                divSum += vmean / nm;
                v2r[V] = r;
                ssp[V + 1] = S + size;
            }
            S += size;
            ++V;
        }
        if (rp) rp[r + 1] = V;
    }
    free(tr);
    free(tm);
    if (out_v) *out_v = V;
    if (out_s) *out_s = S;
    if (div) *div = V ? divSum / V : 0.0;
    return 0;
}

/* ------------------------------------------------------- BVSS BFS (level replay) */
static inline void or32(uint32_t* w, uint32_t bits) { __atomic_fetch_or(w, bits, __ATOMIC_RELAXED); }

uint32_t oracle_bvss_bfs(uint32_t n, uint32_t V, const uint64_t* ssp, const uint32_t* v2r, const uint32_t* rp,
                         const uint32_t* rowIds, const uint32_t* masks, const uint32_t* csrPtr,
                         const uint32_t* csrIdx, uint32_t src, int allowBU, const int32_t* schedule,
                         uint32_t scheduleLen, int32_t* levels, oracle_level* perLevel, uint32_t maxLevels,
                         int threads) {
    if (threads <= 0) threads = omp_get_max_threads();
    const uint32_t W = (n + 31) / 32; /* 32-bit words, as the reference's noWords */
    uint32_t* visited = (uint32_t*)calloc(W ? W : 1, 4);
    uint32_t* next = (uint32_t*)calloc(W ? W : 1, 4);
    uint32_t* front = (uint32_t*)calloc(W ? W : 1, 4);
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
    uint32_t* list2 = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
    const uint8_t* fbyte = (const uint8_t*)front;
    for (uint32_t v = 0; v < n; ++v) levels[v] = -1;
    levels[src] = 0;
    visited[src >> 5] = next[src >> 5] = front[src >> 5] = 1u << (src & 31);
    uint32_t cnt = 0;
    for (uint32_t v = rp[src / 8]; v < rp[src / 8 + 1]; ++v) list[cnt++] = v;
    uint64_t unvisitedLag = (uint64_t)n - 1; /* BFS/BVSSBFSKernel.cuh:1438 */
    uint32_t level = 0;
    int anyNew = 1;
    while (anyNew) {
        ++level;
        int bu;
        if (schedule && level <= scheduleLen) bu = schedule[level - 1] != 0;
        else if (schedule) bu = 0;
        else bu = allowBU && unvisitedLag < (uint64_t)cnt * 10; /* BFS/BVSSBFSKernel.cuh:96 */
        if (!csrPtr) bu = 0;
        uint64_t slices = 0, unv = 0, scanned = 0;
        if (bu) {
            /* BFS/BVSSBFSKernel.cuh:98-129: pull over sorted in-neighbours, stop at the
             * first frontier member. */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024) reduction(+ : unv, scanned)
            for (uint32_t u = 0; u < n; ++u) {
                if ((visited[u >> 5] >> (u & 31)) & 1u) continue;
                ++unv;
                for (uint32_t p = csrPtr[u]; p < csrPtr[u + 1]; ++p) {
                    ++scanned;
                    uint32_t v = csrIdx[p];
                    if ((front[v >> 5] >> (v & 31)) & 1u) {
                        or32(&next[u >> 5], 1u << (u & 31));
                        break;
                    }
                }
            }
        } else {
            /* BFS/BVSSBFSKernel.cuh:133-203: a slice hits when its mask byte shares a
             * bit with the frontier byte of its rset. */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256) reduction(+ : slices)
            for (uint32_t i = 0; i < cnt; ++i) {
                uint32_t vset = list[i];
                uint32_t f = fbyte[v2r[vset]];
                slices += ssp[vset + 1] - ssp[vset];
                for (uint64_t k = ssp[vset]; k < ssp[vset + 1]; ++k) {
                    uint32_t m = (masks[k >> 2] >> (8 * (k & 3))) & 0xFFu;
                    if (m & f) {
                        uint32_t row = rowIds[k];
                        or32(&next[row >> 5], 1u << (row & 31));
                    }
                }
            }
        }
        /* Dense update (BFS/BVSSBFSKernel.cuh:211-262). */
        uint64_t unvNow = 0, newV = 0;
        uint32_t cnt2 = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : unvNow, newV)
        for (uint32_t i = 0; i < W; ++i) {
            uint32_t cur = visited[i];
            unvNow += (uint64_t)__builtin_popcount(~cur);
            uint32_t nx = next[i];
            uint32_t diff = cur ^ nx;
            front[i] = diff;
            if (diff) {
                visited[i] = nx;
                newV += (uint64_t)__builtin_popcount(diff);
                for (uint32_t b = diff; b; b &= b - 1) levels[i * 32 + __builtin_ctz(b)] = (int32_t)level;
                for (uint32_t set = 0; set < 4; ++set)
                    if ((diff >> (8 * set)) & 0xFFu) {
                        uint32_t r = i * 4 + set;
                        uint32_t k = rp[r + 1] - rp[r];
                        uint32_t pos = __atomic_fetch_add(&cnt2, k, __ATOMIC_RELAXED);
                        for (uint32_t v = rp[r]; v < rp[r + 1]; ++v) list2[pos++] = v;
                    }
            }
        }
        if (perLevel && level <= maxLevels) {
            oracle_level* L = &perLevel[level - 1];
            L->bottom_up = bu;
            L->frontier_vsets = bu ? 0 : cnt;
            L->new_vertices = (uint32_t)newV;
            L->next_vsets = cnt2;
            L->slices = slices;
            L->unvisited = unv;
            L->scanned = scanned;
            /* Algorithmic bytes of the reference formulation:
             *   B_td = sum_v (5 s_v + 25) + 12 W + 4 |F'| + 4 |VF'|
             *   B_bu = 4 W + n/8 + sum_u (8 + 4 e_u) + 12 W + 4 |F'|            */
            if (bu)
                L->bytes = 4ull * W + ((uint64_t)n + 7) / 8 + 8ull * unv + 4ull * scanned + 12ull * W + 4ull * newV;
            else
                L->bytes = 5ull * slices + 25ull * cnt + 12ull * W + 4ull * newV + 4ull * cnt2;
        }
        unvisitedLag = unvNow;
        uint32_t* t = list;
        list = list2;
        list2 = t;
        cnt = cnt2;
        anyNew = newV != 0;
    }
    free(visited);
    free(next);
    free(front);
    free(list);
    free(list2);
    /* levels processed = max distance + 1 (BFS/BVSSBFSKernel.cuh:1551-1561) */
    return level;
}

/* --------------------------------------------------------- label propagation CC */
uint32_t oracle_bvss_cc(uint32_t n, uint32_t V, const uint64_t* ssp, const uint32_t* v2r, const uint32_t* rp,
                        const uint32_t* rowIds, const uint32_t* masks, uint32_t* labels) {
    const uint32_t W = (n + 31) / 32;
    uint32_t* front = (uint32_t*)malloc(4 * (size_t)(W ? W : 1));
    uint32_t* marker = (uint32_t*)calloc(W ? W : 1, 4);
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
    uint32_t* list2 = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
    const uint8_t* fbyte = (const uint8_t*)front;
    for (uint32_t v = 0; v < n; ++v) labels[v] = v; /* CC/BVSSCCKernel.cuh:1084-1088 */
    for (uint32_t i = 0; i < W; ++i) front[i] = 0xFFFFFFFFu;
    if (n & 31) front[W - 1] = (1u << (n & 31)) - 1; /* :1183-1188 */
    uint32_t cnt = V;
    for (uint32_t v = 0; v < V; ++v) list[v] = v;
    uint32_t it = 0;
    while (cnt) {
        ++it;
        for (uint32_t i = 0; i < cnt; ++i) { /* :67-221 */
            uint32_t vset = list[i], r = v2r[vset], f = fbyte[r];
            for (uint64_t k = ssp[vset]; k < ssp[vset + 1]; ++k) {
                uint32_t valid = (masks[k >> 2] >> (8 * (k & 3))) & f & 0xFFu;
                if (!valid) continue;
                uint32_t row = rowIds[k], best = labels[row];
                int updated = 0;
                for (; valid; valid &= valid - 1) {
                    uint32_t nb = labels[r * 8 + __builtin_ctz(valid)];
                    if (nb < best) {
                        best = nb;
                        updated = 1;
                    }
                }
                if (updated && best < labels[row]) {
                    labels[row] = best;
                    marker[row >> 5] |= 1u << (row & 31);
                }
            }
        }
        uint32_t c2 = 0; /* :228-269 */
        for (uint32_t i = 0; i < W; ++i) {
            uint32_t mk = marker[i];
            front[i] = mk;
            if (mk)
                for (uint32_t set = 0; set < 4; ++set)
                    if ((mk >> (8 * set)) & 0xFFu) {
                        uint32_t r = i * 4 + set;
                        for (uint32_t v = rp[r]; v < rp[r + 1]; ++v) list2[c2++] = v;
                    }
            marker[i] = 0;
        }
        uint32_t* t = list;
        list = list2;
        list2 = t;
        cnt = c2;
    }
    free(front);
    free(marker);
    free(list);
    free(list2);
    return it;
}

/* ------------------------------------------------------------------ Jaccard window */
/* Plain restatement of CSC::jackardWithWindow (DataStructures/CSC.cuh:717-920), one
 * window after another (the reference runs windows in parallel; they are independent).
 * Per slot of slice_size columns inside a window: the seed is the unplaced column of
 * largest degree, the LAST such index on ties (`>=`, CSC.cuh:772-779); every further
 * column is the queued candidate (an unplaced window column sharing a row with the
 * slot's union) of largest inter / (union + deg - inter), the FIRST in queue order on
 * ties (`>`, CSC.cuh:853), taken out by moving the queue's last entry into its place
 * (CSC.cuh:860-862); with an empty queue the seed rule picks again.  A row's scan of
 * the transpose starts at its first column >= the window start and stops at the first
 * column >= the window end (CSC.cuh:795-826).  inv_perm[old] = new. */
void oracle_jaccard_window(uint32_t n, const uint32_t* cp, const uint32_t* rows, uint32_t slice, uint32_t window,
                           uint32_t* inv_perm) {
    /* transpose: tp / ti list, for each row r, the columns c with r in column c (sorted) */
    uint64_t nnz = cp[n];
    uint32_t* tp = calloc((size_t)n + 1, sizeof(uint32_t));
    uint32_t* ti = malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    for (uint64_t e = 0; e < nnz; ++e) tp[rows[e] + 1]++;
    for (uint32_t r = 0; r < n; ++r) tp[r + 1] += tp[r];
    uint32_t* fill = malloc(((size_t)n + 1) * sizeof(uint32_t));
    memcpy(fill, tp, ((size_t)n + 1) * sizeof(uint32_t));
    for (uint32_t c = 0; c < n; ++c)
        for (uint32_t e = cp[c]; e < cp[c + 1]; ++e) ti[fill[rows[e]]++] = c;
    free(fill);
    const uint32_t NONE = 0xFFFFFFFFu;
    unsigned char* placed = malloc(window ? window : 1);
    uint32_t* rowStart = malloc(((size_t)n ? n : 1) * sizeof(uint32_t));
    uint32_t* inter = malloc((window ? window : 1) * sizeof(uint32_t));
    uint32_t* queue = malloc((window ? window : 1) * sizeof(uint32_t));
    uint32_t* rowMark = calloc((size_t)n ? n : 1, sizeof(uint32_t));
    uint32_t epoch = 1;
    for (uint32_t ws = 0; ws < n; ws += window) {
        const uint32_t we = ws + window < n ? ws + window : n;
        memset(placed, 0, window);
        for (uint32_t r = 0; r < n; ++r) rowStart[r] = NONE;
        for (uint32_t ss = ws; ss < we; ss += slice) {
            const uint32_t se = ss + slice < we ? ss + slice : we;
            memset(inter, 0, window * sizeof(uint32_t));
            uint32_t qn = 0, uni = 0;
            ++epoch;
            for (uint32_t newCol = ss; newCol < se; ++newCol) {
                uint32_t col = NONE;
                if (qn == 0) { /* seed: unplaced column of largest degree, last on ties */
                    uint32_t best = 0;
                    for (uint32_t j = ws; j < we; ++j)
                        if (!placed[j - ws] && cp[j + 1] - cp[j] >= best) {
                            col = j;
                            best = cp[j + 1] - cp[j];
                        }
                    if (col == NONE) break;
                } else { /* best Jaccard ratio in queue order, first on ties */
                    uint32_t bi = 0;
                    double bv = -1.0;
                    for (uint32_t q = 0; q < qn; ++q) {
                        const uint32_t j = queue[q];
                        const uint32_t in = inter[j - ws], deg = cp[j + 1] - cp[j];
                        const uint32_t un = uni + deg - in;
                        const double v = un == 0 ? 1.0 : (double)in / un;
                        if (v > bv) {
                            bv = v;
                            bi = q;
                        }
                    }
                    col = queue[bi];
                    queue[bi] = queue[qn - 1];
                    --qn;
                }
                placed[col - ws] = 1;
                /* absorb the column's rows that are new to the union */
                for (uint32_t e = cp[col]; e < cp[col + 1]; ++e) {
                    const uint32_t r = rows[e];
                    if (rowMark[r] == epoch) continue;
                    rowMark[r] = epoch;
                    ++uni;
                    if (rowStart[r] == NONE) {
                        for (uint32_t p = tp[r]; p < tp[r + 1]; ++p)
                            if (ti[p] >= ws) {
                                rowStart[r] = p;
                                break;
                            }
                        if (rowStart[r] == NONE) continue;
                    }
                    for (uint32_t p = rowStart[r]; p < tp[r + 1]; ++p) {
                        const uint32_t c = ti[p];
                        if (c >= we) break;
                        if (!placed[c - ws]) {
                            if (inter[c - ws] == 0) queue[qn++] = c;
                            inter[c - ws]++;
                        }
                    }
                }
                inv_perm[col] = newCol;
            }
        }
    }
    free(tp);
    free(ti);
    free(placed);
    free(rowStart);
    free(inter);
    free(queue);
    free(rowMark);
}
