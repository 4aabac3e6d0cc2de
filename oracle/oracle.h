/*
 * CPU oracle for the V:N:M reorder + SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded restatement of the
 * reference algorithm (PictureResearch/vnm-graph-reordering) used to check the
 * HIP product library bit-for-bit and to time the CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  All integer
 * outputs (counts, scores, permutations, stats) are defined to be identical
 * to the reference on the same inputs; the quirk decisions are listed in
 * oracle.cpp next to the code that implements them.
 */
#ifndef VNM_ORACLE_H
#define VNM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host COO in the reference's Mtx<float> host layout (mtx/Mtx.h:64-65):
 * entries in COO order, 0-based ids, nrows/ncols already padded to M. */
typedef struct {
    int32_t nrows;
    int32_t ncols;
    int64_t nnz;
    int32_t *row;
    int32_t *col;
    float   *val;
} orc_coo;

/* Status codes (shared numbering with include/vnm.h). */
#define ORC_OK         0
#define ORC_E_IO      -1
#define ORC_E_FORMAT  -2
#define ORC_E_ARG     -3
#define ORC_E_OOM     -5
#define ORC_E_UB_PATH -6

/* ---- MatrixMarket I/O (mtx/readMtx.hpp:354-425, mtx/Mtx.h:98-152,412-435) */
int  orc_mtx_read(const char *path, int M, orc_coo *out);
int  orc_mtx_write(const char *path, const orc_coo *coo);
/* Post-load processing of raw tuples as readMtx does after readTuples:
 * symmetric mirroring, sort, dedup (with the value quirk), sort, M-padding. */
int  orc_coo_from_tuples(int nrows, int ncols, int symmetric, int64_t n,
                         const int32_t *row, const int32_t *col,
                         const float *val, int M, orc_coo *out);
void orc_coo_free(orc_coo *coo);

/* ---- N:M primitives (reorder/reorder.h) */
/* cnt[seg] = #rows with popc(bitmap) > N; zero[seg] = #rows with bitmap 0.
 * Either output may be NULL.  nseg = ceil(nrows / M). */
int orc_locate(const orc_coo *coo, int N, int M, int32_t *cnt, int32_t *zero);
/* score[(b1*M+b2)*2 + {0,1}] = {lgain, rgain} for seg1 x seg2. */
int orc_swap_score(const orc_coo *coo, int N, int M, int seg1, int seg2,
                   int32_t *score);
/* Pop order of the reference SwapPairList built from one score vector:
 * order[k] = index i (0..M*M-1) of the k-th popped SwapPair. */
int orc_pq_pop_order(const int32_t *score, int M, int32_t *order);

/* One swap decision of the greedy, recorded for the trace. */
typedef struct {
    int32_t round;
    int32_t primary;
    int32_t target;
    int32_t v1, v2;
    int32_t lgain, rgain;
    int32_t primary_cnt_after;
    int32_t target_cnt_after;
} orc_swap_rec;

/* Greedy N:M reorder: sched 0..3 = run_onepair[_double|_reverse|_reverse_double]
 * (reorder/reorder.h:1805-2474), sched >= 8 = run (reorder/reorder.h:963-1100).
 * Relabels coo->row/col in place.  stats = {init_size, final_size, init_cnt,
 * final_cnt, rounds}.  trace/trace_cap/trace_len are optional (NULL/0). */
int orc_nm_reorder(orc_coo *coo, int N, int M, int sched, int maxiter,
                   int32_t *stats, orc_swap_rec *trace, int64_t trace_cap,
                   int64_t *trace_len);

/* Allpairs schedules 4..7, M = 4 (reorder/reorder.h:1102-1803): clean = 0
 * is the reference as written, clean = 1 filters every move by visited
 * nodes and applies the double swap that was scored. */
int orc_allpairs_reorder(orc_coo *coo, int N, int sched, int maxiter, int clean,
                         int32_t *stats, int64_t *overwrites);
/* GetSegSwapScoreAll (reorder/reorder.h:133-260): 52 {lgain, rgain}. */
int orc_swap_score_all(const orc_coo *coo, int N, int seg1, int seg2, int32_t *score);

/* ---- VENOM (reordering.py) */
int orc_venom_eval(const orc_coo *coo, int M, int V, int32_t *inv,
                   int32_t *inv_strict);
/* stats = {init_inv, init_inv_strict, final_inv, final_inv_strict, rounds};
 * relabels positions in place; when the reorderer reverts (worse step) the
 * positions are the previous ones, as the script writes them. */
int orc_venom_reorder(orc_coo *coo, int M, int V, int maxiter, int32_t *stats);

/* The C++ reload of temp.mtx after the VENOM script (main.cu:268). */
int orc_reload(orc_coo *coo, int M);

/* ---- Hybrid driver (main.cu:195-337) */
typedef struct {
    float   nrows, nnz, density;
    float   init_size, final_size, init_cnt, final_cnt, improve_rate;
    int32_t init_mb, final_mb, init_vnm, final_vnm;
    int32_t nm_rounds, vnm_rounds, total_rounds;
} orc_hybrid_stats;
/* Runs the hybrid loop in memory on *coo (already loaded), emulating the
 * temp.mtx round trips.  On return *coo holds the final matrix in the order
 * the reference writes --outmtxfile. */
/* Host threads of the reorder path (1: the reference's sequential loops;
 * more: OpenMP over its row and segment loops, same results). */
void orc_set_threads(int threads);
int orc_hybrid(orc_coo *coo, int N, int M, int V, int rounds, int sched,
               orc_hybrid_stats *st);
/* find_best_2toX (main.cu:72-114) over schedules 0..nsched-1 on copies of
 * *coo; ties on improve_rate go to the lowest schedule index. */
int orc_find_best_sched(const orc_coo *coo, int N, int M, int nsched,
                        int32_t *best);

/* ---- Retained nnz and V:N:M packing semantics */
int orc_retained_nnz(const orc_coo *coo, int V, int N, int M, int64_t *out);
/* C[R x ncolsB] (row-major, ld = ncolsB) = A_vnm * B in FP64, where A_vnm is
 * A pruned to the V:N:M pattern (4 selected columns per VxM block, 2 per row
 * inside them: the 2 largest |v|, ties to the lower column). */
int orc_spmm_vnm_f64(const orc_coo *coo, int V, int M, const float *B,
                     int ncolsB, double *C);
/* The same on `threads` OpenMP threads (the all-cores CPU baseline); the
 * result is bit-identical to orc_spmm_vnm_f64. */
int orc_spmm_vnm_f64_mt(const orc_coo *coo, int V, int M, const float *B, int ncolsB, double *C,
                        int threads);

/* Value quantisation of the reference's text round trips. */
float orc_quant_cpp_write(float v);   /* "%.2f" then fscanf %f */
float orc_quant_py_roundtrip(float v);/* "%.2f" -> float64 -> "%.1e" -> %f */

/* ---- Learned pair proposer (lp_oracle.cpp; an addition of this project,
 * no reference counterpart).  Features X = nrows x 16 (FP64); the loss of the
 * pair head against targets and its gradient (22784 parameters, FP64). */
#define ORC_LP_NPARAMS 22784
int orc_lp_features(const orc_coo *coo, int N, int M, double *X);
int orc_lp_loss_grad(const orc_coo *coo, int N, int M, const double *params,
                     const int32_t *pairs, const double *target, int count,
                     double *loss, double *grad, double *head);

#ifdef __cplusplus
}
#endif
#endif
