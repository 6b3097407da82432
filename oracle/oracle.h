/*
 * oracle.h — CPU restatement of the reference traversal path, used ONLY as test
 * infrastructure: by tests/, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline leg, always as the checker or the baseline, never as the product.
 *
 * Graph convention as the reference CSC (DataStructures/CSC.cuh:118-131): column j
 * lists the rows i of the edges j -> i, columns sorted, no duplicates.
 */
#ifndef BLEST_ORACLE_H
#define BLEST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Textbook FIFO BFS from src along j -> i.  levels[n]: distance, -1 unreachable.
 * Returns the number of vertices reached. */
uint32_t oracle_bfs(uint32_t n, const uint32_t* col_ptrs, const uint32_t* rows, uint32_t src,
                    int32_t* levels);

/* Union-find over every stored edge (undirected view).  labels[v] = smallest vertex ID
 * of v's component.  Returns the number of components. */
uint32_t oracle_cc(uint32_t n, const uint32_t* col_ptrs, const uint32_t* rows, uint32_t* labels);

/* BVSS construction (BVSS.cuh:185-467).  Two-phase: call with slice_set_ptrs == NULL to
 * get *out_v and *out_s; then call again with arrays of V+1, V, R+1, S, S/no_masks.
 * Returns 0, or -1 on an invalid slice size / noMasks (BVSS.cuh:103-110). */
int oracle_bvss_build(uint32_t n, const uint32_t* col_ptrs, const uint32_t* rows, uint32_t slice_size,
                      uint32_t no_masks, int full_padding, uint32_t* out_v, uint64_t* out_s,
                      uint64_t* slice_set_ptrs, uint32_t* virtual_to_real, uint32_t* real_ptrs,
                      uint32_t* row_ids, uint32_t* masks, double* update_divergence);

/* Level-synchronous BFS over the BVSS (slice size 8, 4 masks) as the reference's
 * persistent kernels do it (BFS/BVSSBFSKernel.cuh:33-284, 531-691): top-down from the
 * vset list, bottom-up over the CSR (csr_ptr/csr_idx = in-neighbours), dense update.
 * schedule: NULL = the reference rule (bottom-up iff allow_bottom_up and
 * unvisited < 10 * |frontier vsets|, BFS/BVSSBFSKernel.cuh:96); otherwise
 * schedule[l-1] = 1 forces level l bottom-up (0 = top-down) for l <= schedule_len.
 * Outputs levels[n]; per level (up to max_levels) the direction and the algorithmic
 * bytes of the reference formulation.  threads: OpenMP threads (0 = all).
 * Returns the number of levels (max distance + 1). */
typedef struct oracle_level {
    int32_t bottom_up;
    uint32_t frontier_vsets;   /* |VF_l| (top-down) */
    uint32_t new_vertices;     /* |F_{l+1}| */
    uint32_t next_vsets;       /* |VF_{l+1}| */
    uint64_t slices;           /* sum of s_v over VF_l (top-down) */
    uint64_t unvisited;        /* |U_l| (bottom-up) */
    uint64_t scanned;          /* sum of e_u over U_l (bottom-up) */
    uint64_t bytes;            /* B_td or B_bu */
} oracle_level;

uint32_t oracle_bvss_bfs(uint32_t n, uint32_t no_vsets, const uint64_t* slice_set_ptrs,
                         const uint32_t* virtual_to_real, const uint32_t* real_ptrs,
                         const uint32_t* row_ids, const uint32_t* masks, const uint32_t* csr_ptr,
                         const uint32_t* csr_idx, uint32_t src, int allow_bottom_up,
                         const int32_t* schedule, uint32_t schedule_len, int32_t* levels,
                         oracle_level* per_level, uint32_t max_levels, int threads);

/* Min-label propagation over the BVSS (CC/BVSSCCKernel.cuh:33-276) to its fixed
 * point; labels[n] in the graph's own numbering.  Returns the iteration count. */
uint32_t oracle_bvss_cc(uint32_t n, uint32_t no_vsets, const uint64_t* slice_set_ptrs,
                        const uint32_t* virtual_to_real, const uint32_t* real_ptrs,
                        const uint32_t* row_ids, const uint32_t* masks, uint32_t* labels);

/* CSC::jackardWithWindow (CSC.cuh:717-920) restated plainly: inv_perm[old] = new for
 * the greedy Jaccard fill of each slot of `slice` columns inside windows of `window`
 * columns (window a multiple of slice). */
void oracle_jaccard_window(uint32_t n, const uint32_t* col_ptrs, const uint32_t* rows, uint32_t slice,
                           uint32_t window, uint32_t* inv_perm);

/* Number of OpenMP threads a parallel region gets with threads = 0. */
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
