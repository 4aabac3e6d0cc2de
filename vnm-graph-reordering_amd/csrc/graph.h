// Host-side context and device-resident graph state.
#pragma once

#include <hip/hip_runtime.h>

#include "vnm_internal.h"

struct vnm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t launches = 0;
};

namespace vnm {

// A device buffer owned by a graph.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count);
    void release();
    ~DBuf() { release(); }
};

}  // namespace vnm

// The matrix lives on the device as the reference's COO (mtx/Mtx.h:78-80):
// entries in the reference's COO order (the order at the last load), ids
// current.  Every reorder relabels these arrays in place, as the Reorder
// kernel does (reorder/reorder.h:23-35).
struct vnm_graph {
    vnm_ctx *ctx = nullptr;
    int32_t nrows = 0, ncols = 0, M = 8;
    int64_t nnz = 0;
    vnm::DBuf<int32_t> row, col;
    vnm::DBuf<float> val;
    // composed relabel since creation: original id -> current id
    vnm::DBuf<int32_t> perm_total;
    // scratch reused across calls
    vnm::DBuf<int32_t> cptr, crow;        // CSC of nonzero entries (N:M run)
    vnm::DBuf<int32_t> perm, inv, cnt, nzr, ws;
    vnm::DBuf<uint32_t> tab;              // per-slot row bitmaps (bytes in words)
    size_t tab_slots = 0;
    int32_t nseg() const { return (nrows + M - 1) / M; }
};

namespace vnm {

int graph_scratch(vnm_graph *g);
int build_csc(vnm_graph *g);
int locate_all(vnm_graph *g, int N, int32_t *d_cnt, int32_t *d_nzr);
int relabel(vnm_graph *g, const int32_t *d_perm);
int nm_reorder(vnm_graph *g, int N, int sched, int maxiter, vnm_nm_stats *st);
int venom_eval(vnm_graph *g, int V, int32_t *inv, int32_t *inv_strict);
int venom_reorder(vnm_graph *g, int V, int maxiter, vnm_venom_stats *st);
int reload_sort(vnm_graph *g, bool quantize);
int graph_clone(vnm_graph *src, vnm_graph **out);
int ensure_tab(vnm_graph *g, size_t slots);

}  // namespace vnm
