// Hybrid N:M / VENOM driver (main.cu:195-337) and schedule auto-selection
// (main.cu:72-114).  The reference round-trips the matrix through ./temp.mtx
// and a Python process every round; here the matrix stays on the device and
// the round trip is reproduced by reload_sort (values through the text
// formats, positions re-sorted as the loader does).
#include "graph.h"
#include "hip_common.h"

#include <chrono>
#include <cstring>

namespace vnm {

int hybrid(vnm_graph *g, int N, int V, int rounds, int sched, vnm_hybrid_stats *st) {
    memset(st, 0, sizeof *st);
    if (g->nnz == 0) return VNM_OK;  // main.cu:204 prints an empty line
    st->nrows = (float)g->nrows;
    st->nnz = (float)g->nnz;
    // nrows * nrows is an int product in the reference (main.cu:322)
    const int32_t sq = (int32_t)((uint32_t)g->nrows * (uint32_t)g->nrows);
    st->density = (float)g->nnz * 1.0f / (float)sq;
    int rc;
    int32_t init_mb, init_vnm;
    // the reference evaluates a stale ./temp.mtx here (main.cu:207); the
    // input matrix is evaluated instead
    if ((rc = venom_eval(g, V, &init_mb, &init_vnm))) return rc;
    VNM_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    auto t0 = std::chrono::steady_clock::now();
    vnm_nm_stats first, last;
    if ((rc = nm_reorder(g, N, sched, 10, &first))) return rc;
    last = first;
    int vnm_round = 0, nm_round = 0;
    for (int i = 0; i < rounds; ++i) {
        vnm_venom_stats vs;
        if ((rc = venom_reorder(g, V, 10, &vs))) return rc;  // reordering.py --m M --v V
        vnm_round += vs.rounds;
        if ((rc = reload_sort(g, true))) return rc;         // new Mtx("temp.mtx")
        if ((rc = nm_reorder(g, N, sched, 10, &last))) return rc;
        nm_round += last.rounds;
    }
    VNM_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    auto t1 = std::chrono::steady_clock::now();
    int32_t final_mb, final_vnm;
    if ((rc = venom_eval(g, V, &final_mb, &final_vnm))) return rc;
    st->init_size = (float)first.init_size;
    st->final_size = (float)last.final_size;
    st->init_cnt = (float)first.init_cnt;
    st->final_cnt = (float)last.final_cnt;
    st->improve_rate =
        first.init_cnt != 0 ? (float)(first.init_cnt - last.final_cnt) * 1.0f / (float)first.init_cnt
                            : 0.0f;
    st->init_mb = init_mb;
    st->final_mb = final_mb;
    st->init_vnm = init_vnm;
    st->final_vnm = final_vnm;
    st->nm_rounds = nm_round;
    st->vnm_rounds = vnm_round;
    st->total_rounds = nm_round + vnm_round;
    st->milliseconds = std::chrono::duration<float, std::milli>(t1 - t0).count();
    return VNM_OK;
}

int find_best_sched(vnm_graph *g, int N, int nsched, int32_t *best) {
    int best_id = 0;
    float max_imprv = 0.0f;
    for (int i = 0; i < nsched; ++i) {
        vnm_graph *c = nullptr;
        int rc = graph_clone(g, &c);
        if (rc) return rc;
        vnm_nm_stats s;
        rc = nm_reorder(c, N, i, 10, &s);
        vnm_graph_destroy(c);
        if (rc) return rc;
        if (g->nnz == 0 || s.init_size == 0) {  // main.cu:85
            *best = 0;
            return VNM_OK;
        }
        const float imp = s.init_cnt != 0 ? (float)(s.init_cnt - s.final_cnt) * 1.0f / (float)s.init_cnt : 0.0f;
        if (imp > max_imprv) {  // ties keep the lower index (the reference compares wall time)
            best_id = i;
            max_imprv = imp;
        }
    }
    *best = best_id;
    return VNM_OK;
}

}  // namespace vnm

using namespace vnm;

extern "C" {

int vnm_hybrid(vnm_graph *g, int N, int V, int rounds, int sched, vnm_hybrid_stats *st) {
    if (!g || !st) return set_error(VNM_E_ARG, "null argument");
    if (V <= 0) return set_error(VNM_E_ARG, "V must be positive (got %d)", V);
    if (rounds < 0) return set_error(VNM_E_ARG, "rounds must be >= 0");
    return hybrid(g, N, V, rounds, sched, st);
}

int vnm_find_best_sched(vnm_graph *g, int N, int nsched, int32_t *best) {
    if (!g || !best || nsched < 1 || nsched > 4)
        return set_error(VNM_E_ARG, "find_best: nsched must be 1..4");
    return find_best_sched(g, N, nsched, best);
}

}  // extern "C"
