// Hybrid N:M / VENOM driver (main.cu:195-337) and schedule auto-selection
// (main.cu:72-114).  The reference round-trips the matrix through ./temp.mtx
// and a Python process every round; here the matrix stays on the device and
// the round trip is reproduced by reload_sort (values through the text
// formats, positions re-sorted as the loader does).
#include "graph.h"
#include "hip_common.h"

#include <chrono>
#include <cstring>
#include <functional>
#include <vector>

namespace vnm {

using NMStep = std::function<int(vnm_graph *, vnm_nm_stats *)>;

static int hybrid_with(vnm_graph *g, int V, int rounds, const NMStep &nm_step, vnm_hybrid_stats *st) {
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
    if ((rc = nm_step(g, &first))) return rc;
    last = first;
    int vnm_round = 0, nm_round = 0;
    for (int i = 0; i < rounds; ++i) {
        vnm_venom_stats vs;
        if ((rc = venom_reorder(g, V, 10, &vs))) return rc;  // reordering.py --m M --v V
        vnm_round += vs.rounds;
        if ((rc = reload_sort(g, true))) return rc;         // new Mtx("temp.mtx")
        if ((rc = nm_step(g, &last))) return rc;
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

int hybrid(vnm_graph *g, int N, int V, int rounds, int sched, vnm_hybrid_stats *st) {
    return hybrid_with(g, V, rounds,
                       [=](vnm_graph *h, vnm_nm_stats *s) { return nm_reorder(h, N, sched, 10, s); }, st);
}

// The hybrid loop with the learned N:M reorder as its N:M step (an addition
// of this library; the VENOM steps and the stats are the reference's).
int hybrid_learned(vnm_graph *g, const vnm_lp_model *m, int N, int V, int rounds, int mode, int topk,
                   int window, int maxiter, vnm_hybrid_stats *st) {
    return hybrid_with(
        g, V, rounds,
        [=](vnm_graph *h, vnm_nm_stats *s) {
            vnm_lp_stats ls;
            int rc = lp_reorder(h, m, N, mode, topk, window, maxiter, 1, &ls);
            if (rc) return rc;
            s->init_size = ls.init_size;
            s->final_size = ls.final_size;
            s->init_cnt = ls.init_cnt;
            s->final_cnt = ls.final_cnt;
            s->rounds = ls.iters;
            return VNM_OK;
        },
        st);
}

int find_best_sched(vnm_graph *g, int N, int nsched, int32_t *best) {
    *best = 0;
    if (g->nnz == 0) return VNM_OK;  // main.cu:85
    if (g->M != 4 && g->M != 8)
        return set_error(VNM_E_ARG, "the GPU N:M engine supports M = 4 or 8 (got %d)", g->M);
    // the schedules run together on fresh copies
    std::vector<vnm_graph *> c(nsched, nullptr);
    struct Free {
        std::vector<vnm_graph *> &c;
        ~Free() {
            for (vnm_graph *x : c) vnm_graph_destroy(x);
        }
    } free_copies{c};
    int rc;
    for (int i = 0; i < nsched; ++i)
        if ((rc = graph_clone(g, &c[i]))) return rc;
    std::vector<int> sched(nsched);
    std::vector<vnm_nm_stats> s(nsched);
    for (int i = 0; i < nsched; ++i) sched[i] = i;
    if ((rc = nm_reorder_batch(c.data(), nsched, N, sched.data(), 10, s.data()))) return rc;
    if (s[0].init_size == 0) return VNM_OK;  // main.cu:85
    int best_id = 0;
    float max_imprv = 0.0f;
    for (int i = 0; i < nsched; ++i) {
        const float imp =
            s[i].init_cnt != 0 ? (float)(s[i].init_cnt - s[i].final_cnt) * 1.0f / (float)s[i].init_cnt : 0.0f;
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

int vnm_hybrid_learned(vnm_graph *g, const vnm_lp_model *m, int N, int V, int rounds, int mode, int topk,
                       int window, int maxiter, vnm_hybrid_stats *st) {
    if (!g || !st) return set_error(VNM_E_ARG, "null argument");
    if (V <= 0) return set_error(VNM_E_ARG, "V must be positive (got %d)", V);
    if (rounds < 0) return set_error(VNM_E_ARG, "rounds must be >= 0");
    return hybrid_learned(g, m, N, V, rounds, mode, topk, window, maxiter, st);
}

int vnm_graph_clone(vnm_graph *g, vnm_graph **out) {
    if (!g || !out) return set_error(VNM_E_ARG, "null argument");
    return graph_clone(g, out);
}

int vnm_find_best_sched(vnm_graph *g, int N, int nsched, int32_t *best) {
    if (!g || !best || nsched < 1 || nsched > 4)
        return set_error(VNM_E_ARG, "find_best: nsched must be 1..4");
    return find_best_sched(g, N, nsched, best);
}

}  // extern "C"
