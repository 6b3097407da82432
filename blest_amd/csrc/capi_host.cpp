// extern "C" entry points for the host-side graph, reordering and BVSS objects.
#include <cstdio>
#include <cstring>
#include <new>

#include "bvss.hpp"
#include "capi_internal.hpp"
#include "reorder.hpp"

struct blest_csc {
    blest::CSC g;
};
struct blest_bvss {
    blest::BVSS b;
};

namespace blest {

thread_local std::string tl_last_error;

blest_status fail(blest_status s, const std::string& msg) {
    tl_last_error = msg;
    return s;
}

const CSC& csc_of(const blest_csc* h) { return h->g; }
const BVSS& bvss_of(const blest_bvss* h) { return h->b; }
blest_bvss* bvss_wrap(BVSS&& b) { return new blest_bvss{std::move(b)}; }

}  // namespace blest

using namespace blest;

extern "C" {

const char* blest_status_str(blest_status s) {
    switch (s) {
        case BLEST_OK: return "ok";
        case BLEST_ERR_INVALID_ARGUMENT: return "invalid argument";
        case BLEST_ERR_SLICE_SIZE: return "invalid slice size / noMasks";
        case BLEST_ERR_PERMUTATION: return "permutation check has failed";
        case BLEST_ERR_IO: return "i/o error";
        case BLEST_ERR_OUT_OF_HOST_MEMORY: return "out of host memory";
        case BLEST_ERR_OUT_OF_DEVICE_MEMORY: return "out of device memory";
        case BLEST_ERR_DEVICE: return "device (HIP) error";
        case BLEST_ERR_TIMEOUT: return "bounded device wait expired";
        case BLEST_ERR_UNSUPPORTED: return "unsupported";
        case BLEST_ERR_BUSY: return "busy";
    }
    return "unknown status";
}

const char* blest_last_error(void) { return tl_last_error.c_str(); }
const char* blest_version(void) { return "blest-mi355x 0.1.0 (gfx950)"; }

blest_status blest_csc_create(uint32_t n, const uint32_t* col_ptrs, const uint32_t* rows, blest_csc** out) {
    return guarded([&] {
        if (!out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "out is null");
        *out = new blest_csc{csc_from_arrays(n, col_ptrs, rows)};
    });
}

blest_status blest_csc_from_mtx(const char* path, int undirected, int binary, blest_csc** out) {
    return guarded([&] {
        if (!path || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *out = new blest_csc{csc_from_mtx(path, undirected != 0, binary != 0)};
    });
}

blest_status blest_csc_load_binary(const char* path, blest_csc** out) {
    return guarded([&] {
        if (!path || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *out = new blest_csc{csc_load_binary(path)};
    });
}

blest_status blest_csc_save_binary(const blest_csc* g, const char* path) {
    return guarded([&] {
        if (!g || !path) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        csc_save_binary(g->g, path);
    });
}

blest_status blest_csc_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                     uint64_t seed, int permute, blest_csc** out) {
    return guarded([&] {
        if (!out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "out is null");
        *out = new blest_csc{generate_rmat(scale, edge_factor, a, b, c, seed, permute != 0)};
    });
}

blest_status blest_csc_generate_grid(uint32_t side, double keep, double shortcut_frac, uint64_t seed,
                                     blest_csc** out) {
    return guarded([&] {
        if (!out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "out is null");
        *out = new blest_csc{generate_grid(side, keep, shortcut_frac, seed)};
    });
}

blest_status blest_csc_generate_er(uint32_t n, double avg_degree, uint64_t seed, blest_csc** out) {
    return guarded([&] {
        if (!out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "out is null");
        *out = new blest_csc{generate_er(n, avg_degree, seed)};
    });
}

void blest_csc_free(blest_csc* g) { delete g; }

blest_status blest_csc_info(const blest_csc* g, uint32_t* n, uint64_t* nnz, int* is_social) {
    return guarded([&] {
        if (!g) throw Error(BLEST_ERR_INVALID_ARGUMENT, "graph is null");
        if (n) *n = g->g.n;
        if (nnz) *nnz = g->g.nnz();
        if (is_social) *is_social = g->g.isSocial ? 1 : 0;
    });
}

blest_status blest_csc_export(const blest_csc* g, uint32_t* col_ptrs, uint32_t* rows) {
    return guarded([&] {
        if (!g) throw Error(BLEST_ERR_INVALID_ARGUMENT, "graph is null");
        if (col_ptrs) std::memcpy(col_ptrs, g->g.colPtrs.data(), 4 * g->g.colPtrs.size());
        if (rows && g->g.nnz()) std::memcpy(rows, g->g.rows.data(), 4 * g->g.nnz());
    });
}

blest_status blest_csc_is_social(const blest_csc* g, int* is_social) {
    return guarded([&] {
        if (!g || !is_social) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *is_social = csc_is_social(g->g) ? 1 : 0;
    });
}

blest_status blest_csc_is_symmetric(const blest_csc* g, int* symmetric) {
    return guarded([&] {
        if (!g || !symmetric) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *symmetric = csc_is_symmetric(g->g) ? 1 : 0;
    });
}

blest_status blest_csc_metrics_get(const blest_csc* g, blest_csc_metrics* m) {
    return guarded([&] {
        if (!g || !m) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        CSCMetrics x = csc_metrics(g->g);
        m->max_bandwidth = x.maxBandwidth;
        m->max_profile = x.maxProfile;
        m->max_degree = x.maxDegree;
        m->_pad = 0;
        m->avg_bandwidth = x.avgBandwidth;
        m->avg_profile = x.avgProfile;
        m->avg_degree = x.avgDegree;
    });
}

blest_status blest_csc_pick_roots(const blest_csc* g, uint32_t count, uint64_t seed, uint32_t* roots) {
    return guarded([&] {
        if (!g || (!roots && count)) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        std::vector<uint32_t> r = pick_roots(g->g, count, seed);
        std::copy(r.begin(), r.end(), roots);
    });
}

void blest_reorder_opts_default(blest_reorder_opts* o) {
    if (!o) return;
    o->kind = BLEST_REORDER_AUTO;
    o->slice_size = 8;
    o->window = 65536;
    o->jaccard = 1;
    o->gorder = 0;
    o->threads = 0;
    o->seed = 1;
}

blest_status blest_csc_reorder(blest_csc* g, const blest_reorder_opts* o, uint32_t* inv_perm) {
    return guarded([&] {
        if (!g || !o) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        std::vector<uint32_t> inv = reorder(g->g, *o);
        if (inv_perm) std::copy(inv.begin(), inv.end(), inv_perm);
    });
}

blest_status blest_csc_apply_permutation(blest_csc* g, const uint32_t* inv_perm) {
    return guarded([&] {
        if (!g) throw Error(BLEST_ERR_INVALID_ARGUMENT, "graph is null");
        csc_apply_permutation(g->g, inv_perm);
    });
}

blest_status blest_ordering_save(const char* path, const uint32_t* inv_perm, uint32_t n) {
    return guarded([&] {
        if (!path || (!inv_perm && n)) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        FILE* f = std::fopen(path, "wb");
        if (!f) throw Error(BLEST_ERR_IO, std::string("cannot write ordering ") + path);
        size_t w = std::fwrite(inv_perm, 4, n, f);
        std::fclose(f);
        if (w != n) throw Error(BLEST_ERR_IO, "short write");
    });
}

blest_status blest_ordering_load(const char* path, uint32_t* inv_perm, uint32_t n) {
    return guarded([&] {
        if (!path || (!inv_perm && n)) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        FILE* f = std::fopen(path, "rb");
        if (!f) throw Error(BLEST_ERR_IO, std::string("cannot open ordering ") + path);
        size_t r = std::fread(inv_perm, 4, n, f);
        std::fclose(f);
        if (r != n) throw Error(BLEST_ERR_IO, "truncated ordering file");
        if (!permutation_is_bijection(inv_perm, n)) throw Error(BLEST_ERR_PERMUTATION, "Permutation check has failed.");
    });
}

blest_status blest_bvss_build(const blest_csc* g, uint32_t slice_size, uint32_t no_masks, int full_padding,
                              blest_bvss** out) {
    return guarded([&] {
        if (!g || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        bool fp = full_padding < 0 ? !g->g.isSocial : full_padding != 0;
        *out = new blest_bvss{bvss_build(g->g, slice_size, no_masks, fp)};
    });
}

void blest_bvss_free(blest_bvss* b) { delete b; }

blest_status blest_bvss_info_get(const blest_bvss* h, blest_bvss_info* i) {
    return guarded([&] {
        if (!h || !i) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        const BVSS& b = h->b;
        i->n = b.n;
        i->slice_size = b.sliceSize;
        i->no_masks = b.noMasks;
        i->no_real_slice_sets = b.noRealSliceSets;
        i->no_virtual_slice_sets = b.noVirtualSliceSets;
        i->full_padding = b.fullPadding ? 1 : 0;
        i->no_slices = b.noSlices;
        i->no_unpadded_slices = b.noUnpaddedSlices;
        i->no_padded_slices = b.noPaddedSlices;
        i->edges_check = b.edgesCheck;
        i->update_divergence = b.updateDivergence;
        i->avg_slices_per_set = b.avgSlices;
        i->std_slices_per_set = b.stdSlices;
        i->min_slices_per_set = b.minSlices;
        i->max_slices_per_set = b.maxSlices;
        i->compression_ratio = b.compressionRatio;
    });
}

blest_status blest_bvss_export(const blest_bvss* h, uint64_t* ssp, uint32_t* v2r, uint32_t* rp, uint32_t* rows,
                               uint32_t* masks) {
    return guarded([&] {
        if (!h) throw Error(BLEST_ERR_INVALID_ARGUMENT, "bvss is null");
        const BVSS& b = h->b;
        if (ssp) std::copy(b.sliceSetPtrs.begin(), b.sliceSetPtrs.end(), ssp);
        if (v2r) std::copy(b.virtualToReal.begin(), b.virtualToReal.end(), v2r);
        if (rp) std::copy(b.realPtrs.begin(), b.realPtrs.end(), rp);
        if (rows) std::copy(b.rowIds.begin(), b.rowIds.end(), rows);
        if (masks) std::copy(b.masks.begin(), b.masks.end(), masks);
    });
}

blest_status blest_bvss_save_binary(const blest_bvss* b, const char* path) {
    return guarded([&] {
        if (!b || !path) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        bvss_save_binary(b->b, path);
    });
}

blest_status blest_bvss_adopt(uint32_t n, uint32_t slice_size, uint32_t no_masks, uint32_t no_real_slice_sets,
                              uint32_t no_virtual_slice_sets, uint64_t no_slices, const uint64_t* slice_set_ptrs,
                              const uint32_t* virtual_to_real, const uint32_t* real_ptrs, const uint32_t* row_ids,
                              const uint32_t* masks, double update_divergence, int full_padding, blest_bvss** out) {
    return guarded([&] {
        if (!out || !slice_set_ptrs || !real_ptrs) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if ((no_virtual_slice_sets && !virtual_to_real) || (no_slices && (!row_ids || !masks)))
            throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        bvss_check_shape(slice_size, no_masks);
        if (no_slices % no_masks) throw Error(BLEST_ERR_INVALID_ARGUMENT, "BVSS arrays: noSlices is not a whole number of mask words");
        BVSS b;
        b.n = n;
        b.sliceSize = slice_size;
        b.noMasks = no_masks;
        b.noRealSliceSets = no_real_slice_sets;
        b.noVirtualSliceSets = no_virtual_slice_sets;
        b.noSlices = no_slices;
        b.updateDivergence = update_divergence;
        b.fullPadding = full_padding != 0;
        b.sliceSetPtrs.assign(slice_set_ptrs, slice_set_ptrs + size_t(no_virtual_slice_sets) + 1);
        b.virtualToReal.assign(virtual_to_real, virtual_to_real + no_virtual_slice_sets);
        b.realPtrs.assign(real_ptrs, real_ptrs + size_t(no_real_slice_sets) + 1);
        b.rowIds.assign(row_ids, row_ids + no_slices);
        b.masks.assign(masks, masks + no_slices / no_masks);
        bvss_validate(b);
        bvss_compute_stats(b);
        *out = new blest_bvss{std::move(b)};
    });
}

blest_status blest_bvss_load_binary(const char* path, blest_bvss** out) {
    return guarded([&] {
        if (!path || !out) throw Error(BLEST_ERR_INVALID_ARGUMENT, "null argument");
        *out = new blest_bvss{bvss_load_binary(path)};
    });
}

static blest_status copy_str(const std::string& s, char* buf, size_t cap) {
    if (!buf || cap == 0) return fail(BLEST_ERR_INVALID_ARGUMENT, "null buffer");
    size_t k = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
    return s.size() < cap ? BLEST_OK : fail(BLEST_ERR_INVALID_ARGUMENT, "buffer too small");
}

blest_status blest_bvss_csv_header(char* buf, size_t cap) { return copy_str(bvss_csv_header(), buf, cap); }

blest_status blest_bvss_csv_row(const blest_bvss* b, char* buf, size_t cap) {
    if (!b) return fail(BLEST_ERR_INVALID_ARGUMENT, "bvss is null");
    return copy_str(bvss_csv_row(b->b), buf, cap);
}

}  // extern "C"
