// Connected components over the HBM-resident BVSS (gfx950).
//
// Reference: BVSSCCKernel (CC/BVSSCCKernel.cuh:33-276 min-label propagation over the
// BVSS, 1067-1311 host side).  The reference's fixed point is, for an undirected
// graph, the smallest (reordered) vertex ID of each component; its host compaction
// (CC/BVSSCCKernel.cuh:1269-1290) is not reproduced, labels are canonicalised to the
// smallest ORIGINAL vertex ID of the component instead.
//
// algo 0 (default): hook + compress.  Every slice (row i, mask m) of rset r is the set
//   of edges {i, 8r+b : bit b of m}; each undirected edge is hooked once (i < 8r+b).
//   Roots are linked larger-under-smaller with a compare-and-swap, so every parent is
//   smaller than its child and the surviving root of a component is its smallest
//   vertex.  One pass over the slices + one compression pass, independent of diameter.
// algo 1: min-label propagation, the reference's iteration (frontier = rsets whose
//   labels dropped), two stream-ordered kernels per iteration.
#include "dgraph.hpp"

namespace blest {
namespace dev {

__device__ __forceinline__ uint32_t ld_parent(const uint32_t* p, uint32_t x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving.  Parents only ever decrease and always name an
// ancestor, so concurrent halving stores keep the forest valid.
__device__ __forceinline__ uint32_t find_root(uint32_t* p, uint32_t x) {
    uint32_t par = ld_parent(p, x);
    while (par != x) {
        const uint32_t gp = ld_parent(p, par);
        if (gp == par) return par;
        __hip_atomic_store(p + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
        par = ld_parent(p, x);
    }
    return x;
}

__device__ __forceinline__ void hook(uint32_t* p, uint32_t a, uint32_t b) {
    for (;;) {
        uint32_t ra = find_root(p, a), rb = find_root(p, b);
        if (ra == rb) return;
        if (ra > rb) {
            const uint32_t t = ra;
            ra = rb;
            rb = t;
        }
        uint32_t expected = rb;
        if (__hip_atomic_compare_exchange_strong(p + rb, &expected, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        a = ra;
        b = expected;  // rb was hooked meanwhile: continue from its new parent
    }
}

__global__ void cc_init_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

__device__ __forceinline__ void hook_slice(uint32_t* p, uint32_t row, uint32_t m, uint32_t colBase) {
    for (; m; m &= m - 1) {
        const uint32_t col = colBase + __builtin_ctz(m);
        if (row < col) hook(p, row, col);
    }
}

// Wave per 64 vsets, load-balanced over their quads (as the BFS expansion).
__global__ void __launch_bounds__(256) cc_hook_kernel(GraphView g, uint32_t* p) {
    const unsigned lane = lane_id();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    for (unsigned base = wave * kWave; base < g.V; base += nwaves * kWave) {
        const unsigned v = base + lane;
        uint32_t q0 = 0, qc = 0, r = 0;
        if (v < g.V) {
            q0 = g.unitQuad[v];
            qc = g.unitQuad[v + 1] - q0;
            r = g.unitRset[v];
        }
        const unsigned incl = wave_inclusive_scan(qc);
        const unsigned excl = incl - qc;
        const unsigned total = __shfl(incl, kWave - 1, kWave);
        for (unsigned t0 = 0; t0 < total; t0 += kWave) {
            const unsigned t = t0 + lane;
            const unsigned owner = wave_owner(excl, min(t, total - 1));
            const uint32_t oq0 = __shfl(q0, owner, kWave);
            const unsigned oex = __shfl(excl, owner, kWave);
            const uint32_t orr = __shfl(r, owner, kWave);
            if (t < total) {
                const uint32_t q = oq0 + (t - oex);
                const uint4 r4 = g.rows[q];
                const uint32_t m = g.masks[q];
                const uint32_t cb = orr * 8;
                hook_slice(p, r4.x, m & 0xFF, cb);
                hook_slice(p, r4.y, (m >> 8) & 0xFF, cb);
                hook_slice(p, r4.z, (m >> 16) & 0xFF, cb);
                hook_slice(p, r4.w, (m >> 24) & 0xFF, cb);
            }
        }
    }
}

// After hooking the roots are final.  Compression must store roots only: a
// path-halving store (an ancestor that is not the root) racing with another thread's
// final store could overwrite it.  Every value a thread can read here (an old parent or
// a root) is an ancestor, so the read-only walk always ends at the root.
__global__ void cc_compress_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t x = i, par = ld_parent(p, x);
        while (par != x) {
            x = par;
            par = ld_parent(p, x);
        }
        if (x != i) __hip_atomic_store(p + i, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void fill_u32_kernel(uint32_t* a, uint32_t v, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// minOrig[root] = smallest original ID in the component.
__global__ void cc_min_original_kernel(const uint32_t* p, const uint32_t* newToOld, uint32_t* minOrig, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicMin(minOrig + p[i], newToOld[i]);
}

// out[old] = minOrig[root(new(old))]
__global__ void cc_relabel_kernel(const uint32_t* p, const uint32_t* oldToNew, const uint32_t* minOrig,
                                  uint32_t* out, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = minOrig[p[oldToNew[i]]];
}

// ------------------------------------------------------------ label propagation
struct LpArgs {
    uint32_t* labels;
    unsigned long long* front0;
    unsigned long long* front1;
    unsigned long long* marker;
    uint32_t* list0;
    uint32_t* list1;
    Ctrl* ctrl;
    SegView sv;
    SegCounters sc;
};

__device__ __forceinline__ void lp_slice(uint32_t* labels, unsigned long long* marker, uint32_t row, uint32_t valid,
                                         uint32_t colBase) {
    if (!valid) return;
    uint32_t best = ld_parent(labels, row);
    const uint32_t mine = best;
    for (; valid; valid &= valid - 1) best = min(best, ld_parent(labels, colBase + __builtin_ctz(valid)));
    if (best < mine) {
        const uint32_t old = atomicMin(labels + row, best);
        if (best < old) atomicOr(&marker[row >> 6], 1ull << (row & 63));
    }
}

// All vertices start active; every vset is in the first list (CC/BVSSCCKernel.cuh:1084-1088,
// 1183-1203).
__global__ void __launch_bounds__(kProducerBlock) cc_lp_init_kernel(GraphView g, LpArgs a) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < g.n; i += nthreads) a.labels[i] = i;
    for (uint32_t w = tid; w < g.W; w += nthreads) {
        const uint32_t cnt = min(64u, g.n - w * 64);
        a.front1[w] = cnt == 64 ? ~0ull : ((1ull << cnt) - 1);
        a.marker[w] = 0ull;
    }
    // every segment lists all of its vsets
    for (uint32_t v = tid; v < g.V; v += nthreads) a.list1[v] = v;
    for (uint32_t s = tid; s < a.sv.segs; s += nthreads)
        a.sc.count[a.sv.segs + s] = a.sv.segBase[s + 1] - a.sv.segBase[s];
    if (tid == 0) a.ctrl->done = 0;
}

// Iteration L: every slice of a listed vset whose mask meets its rset's active byte
// lowers its row's label to the smallest active column label (CC/BVSSCCKernel.cuh:67-221).
__global__ void __launch_bounds__(kProducerBlock) cc_lp_expand_kernel(GraphView g, LpArgs a, unsigned L) {
    __shared__ unsigned pre[kMaxSegs];
    __shared__ unsigned tmp[16];
    Ctrl* c = a.ctrl;
    if (ld_relaxed(&c->done)) return;
    const unsigned cnt = load_prefix(a.sc.count + size_t(L % 3) * a.sv.segs, a.sv.segs, pre, tmp);
    if (cnt == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->iterations = L - 1;
            __hip_atomic_store(&c->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    const unsigned lane = lane_id();
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = tid / kWave, nwaves = (gridDim.x * blockDim.x) / kWave;
    const unsigned long long* fin = (L & 1) ? a.front1 : a.front0;
    const uint32_t* lin = (L & 1) ? a.list1 : a.list0;
    for (unsigned base = wave * kWave; base < cnt; base += nwaves * kWave) {
        const unsigned e = base + lane;
        uint32_t q0 = 0, qc = 0, fb = 0, r = 0;
        if (e < cnt) {
            const uint32_t v = list_entry(lin, a.sv, pre, e);
            q0 = g.unitQuad[v];
            qc = g.unitQuad[v + 1] - q0;
            r = g.unitRset[v];
            fb = reinterpret_cast<const uint8_t*>(fin)[r];
            if (!fb) qc = 0;
        }
        const unsigned incl = wave_inclusive_scan(qc);
        const unsigned excl = incl - qc;
        const unsigned total = __shfl(incl, kWave - 1, kWave);
        for (unsigned t0 = 0; t0 < total; t0 += kWave) {
            const unsigned t = t0 + lane;
            const unsigned owner = wave_owner(excl, min(t, total - 1));
            const uint32_t oq0 = __shfl(q0, owner, kWave);
            const unsigned oex = __shfl(excl, owner, kWave);
            const uint32_t of = __shfl(fb, owner, kWave);
            const uint32_t orr = __shfl(r, owner, kWave);
            if (t < total) {
                const uint32_t q = oq0 + (t - oex);
                const uint4 r4 = g.rows[q];
                const uint32_t m = g.masks[q];
                const uint32_t cb = orr * 8;
                lp_slice(a.labels, a.marker, r4.x, m & of & 0xFF, cb);
                lp_slice(a.labels, a.marker, r4.y, (m >> 8) & of & 0xFF, cb);
                lp_slice(a.labels, a.marker, r4.z, (m >> 16) & of & 0xFF, cb);
                lp_slice(a.labels, a.marker, r4.w, (m >> 24) & of & 0xFF, cb);
            }
        }
    }
}

// Marked vertices become the next active set; their rsets' vsets the next list
// (CC/BVSSCCKernel.cuh:228-269).
__global__ void __launch_bounds__(kProducerBlock) cc_lp_update_kernel(GraphView g, LpArgs a, unsigned L) {
    __shared__ unsigned tmp[16];
    Ctrl* c = a.ctrl;
    if (ld_relaxed(&c->done)) return;
    unsigned long long* fout = (L & 1) ? a.front0 : a.front1;
    uint32_t* lout = (L & 1) ? a.list0 : a.list1;
    const uint32_t ws = blockIdx.x * a.sv.wps, we = min(g.W, ws + a.sv.wps);
    uint32_t* out = lout + a.sv.segBase[blockIdx.x];
    unsigned running = 0;
    for (uint32_t r0 = ws; r0 < we; r0 += blockDim.x) {
        const uint32_t w = r0 + threadIdx.x;
        unsigned long long mk = 0;
        if (w < we) {
            mk = a.marker[w];
            if (mk) a.marker[w] = 0ull;
            fout[w] = mk;
        }
        const unsigned cnt = word_vsets(g.rsetUnit, w, mk);
        unsigned tot;
        const unsigned ex = block_exclusive_scan(cnt, tmp, tot);
        // each wave writes its 64 words' vsets, balanced over its lanes
        const unsigned waveStart = __shfl(ex, 0, kWave);
        const uint32_t w0 = r0 + (threadIdx.x / kWave) * kWave;
        const uint32_t w1 = min(we, w0 + kWave);
        if (w0 < w1)
            wave_write_vsets(g.rsetUnit, w0, w1, [&](uint32_t x) { return __shfl(mk, x - w0, kWave); },
                             out + running + waveStart);
        running += tot;
    }
    if (threadIdx.x == 0) a.sc.count[size_t((L + 1) % 3) * a.sv.segs + blockIdx.x] = running;
}

}  // namespace dev

// ------------------------------------------------------------------------ host
static void cc_run(Query& q, int algo, uint32_t* hostLabels, blest_cc_stats* st) {
    DeviceGraph& g = *q.g;
    if (!g.symmetric) throw Error(BLEST_ERR_UNSUPPORTED, "connected components need a symmetric graph");
    const bool labelsOnDevice = hostLabels && check_result_range(hostLabels, 4 * size_t(g.n), g.device, "labels");
    if (!q.labels) {
        q.labels = q.alloc<uint32_t>(g.n);
        q.labelsAux = q.alloc<uint32_t>(g.n);
    }
    GraphView gv = g.view();
    const dim3 grid(uint32_t(g.cus) * 8), block(256);
    unsigned iterations = 1;
    BLEST_HIP(hipEventRecord(q.ev0, q.stream));
    if (algo == 0) {
        hipLaunchKernelGGL(dev::cc_init_kernel, grid, block, 0, q.stream, q.labels, g.n);
        hipLaunchKernelGGL(dev::cc_hook_kernel, grid, block, 0, q.stream, gv, q.labels);
        hipLaunchKernelGGL(dev::cc_compress_kernel, grid, block, 0, q.stream, q.labels, g.n);
        BLEST_HIP(hipGetLastError());
    } else if (algo == 1) {
        dev::LpArgs a;
        a.labels = q.labels;
        a.front0 = q.front0;
        a.front1 = q.front1;
        a.marker = q.next;
        a.list0 = q.list0;
        a.list1 = q.list1;
        a.ctrl = q.ctrl;
        a.sv = g.seg_view();
        a.sc = q.seg_counters();
        const dim3 lg = level_grid(g), lb(dev::kProducerBlock);
        hipLaunchKernelGGL(dev::cc_lp_init_kernel, lg, lb, 0, q.stream, gv, a);
        run_levels(q, [&](unsigned L) {
            hipLaunchKernelGGL(dev::cc_lp_expand_kernel, lg, lb, 0, q.stream, gv, a, L);
            hipLaunchKernelGGL(dev::cc_lp_update_kernel, lg, lb, 0, q.stream, gv, a, L);
        });
        BLEST_HIP(hipGetLastError());
    } else {
        throw Error(BLEST_ERR_INVALID_ARGUMENT, "unknown CC algorithm");
    }
    BLEST_HIP(hipEventRecord(q.ev1, q.stream));
    BLEST_HIP(hipStreamSynchronize(q.stream));
    if (algo == 1) {
        dev::Ctrl c;
        BLEST_HIP(hipMemcpyAsync(&c, q.ctrl, sizeof(c), hipMemcpyDeviceToHost, q.stream));
    BLEST_HIP(hipStreamSynchronize(q.stream));
        iterations = c.iterations;
    }
    float ms = 0;
    BLEST_HIP(hipEventElapsedTime(&ms, q.ev0, q.ev1));
    const uint32_t* result = q.labels;
    if (g.hasPerm) {
        hipLaunchKernelGGL(dev::fill_u32_kernel, grid, block, 0, q.stream, q.labelsAux, 0xFFFFFFFFu, g.n);
        hipLaunchKernelGGL(dev::cc_min_original_kernel, grid, block, 0, q.stream, q.labels, g.permNewToOld,
                           q.labelsAux, g.n);
        hipLaunchKernelGGL(dev::cc_relabel_kernel, grid, block, 0, q.stream, q.labels, g.permOldToNew, q.labelsAux,
                           reinterpret_cast<uint32_t*>(q.levels), g.n);
        BLEST_HIP(hipGetLastError());
        result = reinterpret_cast<const uint32_t*>(q.levels);
    }
    // Components = vertices that are their own label, counted on the host from host
    // memory only: when the caller's array is device memory the count reads a host copy.
    if (hostLabels) BLEST_HIP(hipMemcpyAsync(hostLabels, result, 4 * size_t(g.n), hipMemcpyDefault, q.stream));
    std::vector<uint32_t> tmp;
    const uint32_t* counted = hostLabels;
    if (!hostLabels || labelsOnDevice) {
        tmp.resize(g.n);
        BLEST_HIP(hipMemcpyAsync(tmp.data(), result, 4 * size_t(g.n), hipMemcpyDeviceToHost, q.stream));
        counted = tmp.data();
    }
    BLEST_HIP(hipStreamSynchronize(q.stream));
    if (st) {
        uint32_t comps = 0;
        for (uint32_t v = 0; v < g.n; ++v) comps += counted[v] == v;
        st->kernel_ms = ms;
        st->no_components = comps;
        st->iterations = iterations;
    }
}

}  // namespace blest

using namespace blest;

extern "C" blest_status blest_cc(blest_query* qh, int algo, uint32_t* labels, blest_cc_stats* st) {
    return guarded([&] {
        if (!qh) throw Error(BLEST_ERR_INVALID_ARGUMENT, "query is null");
        Query& q = qh->q;
        std::lock_guard<std::mutex> lk(q.mu);
        BLEST_HIP(hipSetDevice(q.g->device));
        cc_run(q, algo, labels, st);
    });
}
