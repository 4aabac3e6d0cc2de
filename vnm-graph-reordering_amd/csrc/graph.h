// Host-side context and device-resident graph state.
#pragma once

#include <hip/hip_runtime.h>

#include "vnm_internal.h"

#include <map>
#include <string>
#include <vector>

// Kernel timers: events recorded around selected launches on the context's
// stream (the reference's GpuTimer, util/utility.h:267-301), resolved lazily.
struct vnm_timer_pending {
    std::string name;
    hipEvent_t a, b;
};

struct vnm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    float *noalloc_slots = nullptr;  // device address of the static SpMM slot array on `device`
    int64_t launches = 0;
    bool timing = false;
    std::vector<vnm_timer_pending> pending;
    std::map<std::string, std::pair<double, int64_t>> timers;
};

namespace vnm {

// A device buffer owned by a graph.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return set_error(e == hipErrorOutOfMemory ? VNM_E_OOM : VNM_E_HIP,
                             "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return VNM_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
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
    vnm::DBuf<uint32_t> rt;               // N:M greedy row records A and B (16 B per row each)
    vnm::DBuf<int32_t> gst;               // N:M greedy device state
    vnm::DBuf<int32_t> queue;             // work counters of the per-segment passes
    // a batch of N:M greedy rounds captured once (its kernels read only the
    // buffers above, which live as long as the graph)
    hipGraph_t nm_graph = nullptr;
    hipGraphExec_t nm_exec = nullptr;
    std::vector<uint8_t> nm_key;          // the arguments nm_exec was captured with
    size_t tab_slots = 0;
    // a relabel of a matrix with fewer columns than rows moved entries past
    // the padded column count (the row permutation also renames columns)
    bool col_overflow = false;
    // bumped whenever the COO is relabelled or re-sorted; caches built from
    // the COO (the proposer's adjacency) are valid for one epoch
    int64_t epoch = 0;
    void *lp_cache = nullptr;
    void (*lp_cache_free)(void *) = nullptr;
    int32_t nseg() const { return (nrows + M - 1) / M; }
    ~vnm_graph() {
        if (lp_cache_free) lp_cache_free(lp_cache);
        if (nm_exec) (void)hipGraphExecDestroy(nm_exec);
        if (nm_graph) (void)hipGraphDestroy(nm_graph);
    }
};

namespace vnm {

int graph_scratch(vnm_graph *g);
int build_csc(vnm_graph *g);
int build_csc_unique(vnm_graph *g);
int locate_all(vnm_graph *g, int N, int32_t *d_cnt, int32_t *d_nzr);
int relabel(vnm_graph *g, const int32_t *d_perm);
int max_ids(vnm_graph *g, const int32_t *row, const int32_t *col, int32_t *max_row, int32_t *max_col);
int nm_reorder(vnm_graph *g, int N, int sched, int maxiter, vnm_nm_stats *st);
int nm_reorder_batch(vnm_graph *const *gs, int count, int N, const int *scheds, int maxiter,
                     vnm_nm_stats *stats);
int nm_reorder_allpairs(vnm_graph *g, int N, int sched, int maxiter, int clean, vnm_nm_stats *stats);
int nm_reorder_ex(vnm_graph *g, int N, int sched, int maxiter, int round_cap, int32_t *d_trace,
                  int64_t trace_cap, int64_t *trace_len, vnm_nm_stats *stats);
int venom_eval(vnm_graph *g, int V, int32_t *inv, int32_t *inv_strict);
int lp_reorder(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window, int maxiter,
               uint64_t seed, vnm_lp_stats *stats);
int lp_reorder_batch(vnm_graph *const *gs, int count, const vnm_lp_model *m, int N, int mode, int topk,
                     int window, int maxiter, uint64_t seed, vnm_lp_stats *stats);
int venom_reorder(vnm_graph *g, int V, int maxiter, vnm_venom_stats *st);
int reload_sort(vnm_graph *g, bool quantize);
int graph_clone(vnm_graph *src, vnm_graph **out);
int ensure_tab(vnm_graph *g, size_t slots);
int noalloc_setup(int device, float **out);

// Event pair around one launch when the context's timers are on.
struct KernelTimer {
    vnm_ctx *ctx;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(vnm_ctx *c, const char *n);
    void stop();
    ~KernelTimer() { stop(); }  // a scope left early still records its end
    KernelTimer(const KernelTimer &) = delete;
    KernelTimer &operator=(const KernelTimer &) = delete;
};

}  // namespace vnm
