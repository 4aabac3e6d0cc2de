/*
 * vnm.h -- C ABI of the MI355X V:N:M reorder + V:N:M SpMM library (libvnm.so).
 *
 * Plain C types only.  Every call returns an int status: VNM_OK (0) or a
 * negative VNM_E_* code; vnm_last_error() gives the message of the last
 * failure on the calling thread.  Nothing is swallowed: HIP failures surface
 * as VNM_E_HIP.  Device work runs on the context's own HIP stream, created
 * non-blocking: it does not wait for work on other streams (torch's current
 * stream included).  Device buffers handed to a call must be complete when
 * the call is made -- synchronize the stream that wrote them, or make the
 * context stream (vnm_ctx_stream) wait on an event recorded after the writes.
 *
 * Each entry point names the reference interface it replaces
 * (PictureResearch/vnm-graph-reordering, file:line).
 */
#ifndef VNM_H
#define VNM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VNM_OK          0
#define VNM_E_IO       -1  /* file cannot be opened/written                */
#define VNM_E_FORMAT   -2  /* bad MatrixMarket banner, size line or entry   */
#define VNM_E_ARG      -3  /* invalid argument (N, M, V, sched, sizes)      */
#define VNM_E_HIP      -4  /* a HIP runtime call or kernel launch failed    */
#define VNM_E_OOM      -5  /* host or device allocation failed             */
#define VNM_E_UB_PATH  -6  /* input reached a path the reference leaves undefined */
#define VNM_E_NODEV    -7  /* no usable GPU                                 */

typedef struct vnm_ctx vnm_ctx;
typedef struct vnm_graph vnm_graph;
typedef struct vnm_packed vnm_packed;

/* Host COO, the reference's Mtx<float> host arrays (mtx/Mtx.h:64-65):
 * 0-based ids, entries in the reference's COO order, dims padded to M. */
typedef struct {
    int32_t nrows;
    int32_t ncols;
    int64_t nnz;
    int32_t *row;
    int32_t *col;
    float   *val;
} vnm_coo;

const char *vnm_last_error(void);

/* ---------------------------------------------------------------- host I/O */

/* Replaces readMtx (mtx/readMtx.hpp:354-425) + Mtx::init_from_file's host part
 * (mtx/Mtx.h:98-112): banner/size parsing, symmetric mirroring, (row, col)
 * sort, duplicate removal (values left in sorted order, as the reference
 * does), padding of both dims to a multiple of M when nrows % M != 0.
 * Out-of-range indices are VNM_E_FORMAT (the reference writes out of bounds).
 * The arrays are allocated by the library; release with vnm_coo_free. */
int vnm_mtx_read(const char *path, int M, vnm_coo *out);

/* vnm_mtx_read with the host worker threads given (0 = the machine's, at
 * most 16; 1 = one thread, the reference's sequential fscanf loop and
 * std::sort, mtx/readMtx.hpp:84-147,56-82).  Any thread count gives the
 * same COO and the same first error. */
int vnm_mtx_read_ex(const char *path, int M, int threads, vnm_coo *out);

/* vnm_mtx_read through a binary cache next to the file, named as the
 * reference's convert() names it (mtx/readMtx.hpp:230-247: ".<file>.<ud|d>.
 * <nosl|sl>.bin"; the reference never writes one): the loader's COO before
 * the M-padding, used while the .mtx keeps its size and modification time,
 * written after a parse when the directory allows.  *from_cache (may be
 * NULL) = 1 when the cache was used.  Same COO as vnm_mtx_read. */
int vnm_mtx_read_cached(const char *path, int M, int threads, int *from_cache, vnm_coo *out);

/* Replaces Mtx::write_to_file (mtx/Mtx.h:412-435): "%%MatrixMarket matrix
 * coordinate real general", "nrows ncols nnz", then "%d %d %.2f" per entry
 * in COO order (1-based). */
int vnm_mtx_write(const char *path, const vnm_coo *coo);

/* vnm_mtx_write with the formatting threads given (0 = the machine's, at
 * most 16; 1 = the reference's single fprintf loop); the bytes written do
 * not depend on it. */
int vnm_mtx_write_ex(const char *path, const vnm_coo *coo, int threads);

/* The loader's ordering applied to raw tuples (mtx/readMtx.hpp:152-219,419,
 * mtx/Mtx.h:105-112). */
int vnm_coo_from_tuples(int32_t nrows, int32_t ncols, int symmetric, int64_t n,
                        const int32_t *row, const int32_t *col, const float *val,
                        int M, vnm_coo *out);
void vnm_coo_free(vnm_coo *coo);

/* Seeded generators (the reference downloads SuiteSparse instead,
 * scripts/tab5_small.sh:920-941).  Output is loader-ordered and padded. */
int vnm_gen_er(int64_t n, int64_t k, uint64_t seed, float lo, float hi, int M, vnm_coo *out);
int vnm_gen_rmat(int scale, int64_t undirected_edges, double a, double b, double c,
                 uint64_t seed, int M, vnm_coo *out);
int vnm_gen_chung_lu(int64_t n, double gamma, double mean_degree, uint64_t seed, int M,
                     vnm_coo *out);
void vnm_gen_normal(uint64_t seed, int64_t count, float *out);

/* ---------------------------------------------------------------- context */

/* One context per host thread: owns a HIP stream on `device` and the
 * device memory of the graphs created on it.  No global or CWD state (the
 * reference shares ./temp.mtx and ./output.txt, main.cu:118-152). */
int  vnm_ctx_create(int device, vnm_ctx **out);
/* Destroy every graph, packed matrix, model and batch created on the
 * context first: their destroy calls wait on the context's stream. */
void vnm_ctx_destroy(vnm_ctx *ctx);
/* The context's hipStream_t, for event timing by the caller. */
void *vnm_ctx_stream(vnm_ctx *ctx);
/* Number of kernels launched on the context so far. */
int64_t vnm_ctx_launches(vnm_ctx *ctx);
/* Kernel timers (the reference's GpuTimer / START_TIMER, util/utility.h:
 * 177-186,267-301): when enabled, HIP events bracket each launch of
 * k_spmm, k_spmm_dense, k_spmm_batch, k_gemm_xw, k_score_rows, k_locate_all
 * and k_relabel, each VENOM iteration ("venom_iteration"), each proposer
 * forward pass ("lp_forward") and pair scoring ("k_lp_pair_gain"), and each
 * whole N:M greedy run ("nm_greedy"), on the context stream.
 * Enabling or disabling resets the totals; reading waits for the events. */
int vnm_ctx_timing(vnm_ctx *ctx, int enable);
int vnm_ctx_timer_read(vnm_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches);

/* ---------------------------------------------------------------- graph */

/* Replaces Mtx::convert_to_bsr_and_sync_device (mtx/Mtx.h:308-379): uploads
 * the COO and builds the device segment stores.  M in {4, 8, 16}.  Positions
 * (row, col) must be distinct, as vnm_mtx_read and vnm_coo_from_tuples give
 * them (the reference's loader removes repeats before csr2bsr,
 * mtx/readMtx.hpp:197-218). */
int  vnm_graph_create(vnm_ctx *ctx, const vnm_coo *coo, int M, vnm_graph **out);
void vnm_graph_destroy(vnm_graph *g);
/* Current COO, in the order the reference writes it (mtx/Mtx.h:429-431):
 * the order at the last (re)load, ids remapped.  Arrays library-allocated. */
int  vnm_graph_download(vnm_graph *g, vnm_coo *out);
/* A device copy of g (current COO and composed relabel) on the same context. */
int  vnm_graph_clone(vnm_graph *g, vnm_graph **out);
/* Map from the ids at vnm_graph_create to current ids (int32[nrows given to
 * vnm_graph_create]).  A reload to the (n, n) shape with n the padded column
 * count (reordering.py:180) can cut off empty rows: their ids stay at or past
 * the current row count. */
int  vnm_graph_perm(vnm_graph *g, int32_t *perm);
/* Current shape and entry count: Mtx's nrows, ncols, nnz (mtx/Mtx.h:54-56),
 * which a VENOM reload resets to the padded column count. */
int  vnm_graph_shape(vnm_graph *g, int32_t *nrows, int32_t *ncols, int64_t *nnz);

/* ---------------------------------------------------------------- N:M primitives */

/* Replaces LocateinvalidSegment / LocateAllZeroSegment + locate_invalid_segment
 * (reorder/reorder.h:262-317,436-496,767-797): cnt[seg] = rows whose bitmap in
 * seg has more than N ones, zero[seg] = rows whose bitmap is 0 (padded rows
 * included).  Host arrays of ceil(nrows/M) ints; either may be NULL. */
int vnm_locate(vnm_graph *g, int N, int32_t *cnt, int32_t *zero);

/* Replaces GetSegSwapScore + get_segment_swap_score (reorder/reorder.h:75-131,
 * 498-535), batched: for k < count, scores[k*2*M*M + (b1*M+b2)*2 + {0,1}] =
 * {lgain, rgain} of exchanging column b1 of seg1 with column b2 of seg2[k]. */
int vnm_swap_score(vnm_graph *g, int N, int seg1, const int32_t *seg2, int count,
                   int32_t *scores);

/* Pop order of the reference SwapPairList (std::priority_queue with the
 * non-strict Comparator, reorder/swappair.h:112-120) for `count` score vectors
 * of 2*M*M ints, computed on the GPU by the same heap emulation the greedy
 * kernel uses: order[k*M*M + j] = index of the j-th popped pair. */
int vnm_pq_pop_order(vnm_ctx *ctx, const int32_t *scores, int count, int M, int32_t *order);

/* ---------------------------------------------------------------- reorder */

typedef struct {
    int32_t init_size, final_size, init_cnt, final_cnt, rounds;
} vnm_nm_stats;

/* Replaces reorderUtil::run_onepair / _double / _reverse / _reverse_double
 * (sched 0..3, reorder/reorder.h:1805-2474) and run (sched >= 8, :963-1100),
 * called as main.cu:19-53 does (maxiter = 10).  The whole greedy runs on the
 * GPU; the stats are the reference's return vector.  sched 4..7 (M = 4) are
 * the allpairs schedules as vnm_nm_reorder_allpairs with clean = 0. */
int vnm_nm_reorder(vnm_graph *g, int N, int sched, int maxiter, vnm_nm_stats *st);

/* Replaces run_allpairs / _double / _reverse / _reverse_double (sched 4..7,
 * M = 4, reorder/reorder.h:1102-1803; unreachable from the reference CLI):
 * single and double column swaps scored by GetSegSwapScoreAll (:133-260).
 * clean = 0 reproduces the reference's visited filter, which never drops
 * single swaps (:1194-1197), and its pairtab double swaps (moves 33..35 apply
 * a different pair than they score, reorder/swappair.h:54-97); when that
 * filter lets a node be swapped twice the reference relabel is no longer a
 * permutation and the call returns VNM_E_UB_PATH.  clean = 1 drops every move
 * touching a swapped node and applies the move that was scored. */
int vnm_nm_reorder_allpairs(vnm_graph *g, int N, int sched, int maxiter, int clean, vnm_nm_stats *st);

/* vnm_nm_reorder on count graphs of one context at once: one chain of round
 * kernels runs every graph's greedy (grid y = graph), so the batch takes
 * about as long as its longest run rather than the sum.  sched[i] is graph
 * i's schedule (0..3 or >= 8), st[i] its stats; each result is that of
 * vnm_nm_reorder on the graph alone.  No counterpart in the reference, which
 * reorders one matrix per process (main.cu:339-352); find_best_2toX's four
 * runs (main.cu:72-114) and batches of small graphs are its callers. */
int vnm_nm_reorder_batch(vnm_graph *const *graphs, int count, int N, const int32_t *sched, int maxiter,
                         vnm_nm_stats *st);

/* vnm_nm_reorder with the reference's swap decisions recorded (what the
 * verbose swap print of run_onepair shows, reorder/reorder.h:1913-1918, as
 * data): 9 ints
 * per swap {round, primary, target, v1, v2, lgain, rgain, primary_cnt,
 * target_cnt}, at most cap records kept, *count = swaps made.  round_cap > 0
 * stops after that many rounds (training traces for the proposer). */
int vnm_nm_trace(vnm_graph *g, int N, int sched, int maxiter, int round_cap, int32_t *rec,
                 int64_t cap, int64_t *count, vnm_nm_stats *st);

/* Replaces `python3 reordering.py --evalonly --m M --v V` (reordering.py:125-198).
 * M (the graph's) 4, 8 or 16; with M = 16 at most 262,112 columns. */
int vnm_venom_eval(vnm_graph *g, int V, int32_t *inv, int32_t *inv_strict);

typedef struct {
    int32_t init_inv, init_inv_strict, final_inv, final_inv_strict, rounds;
} vnm_venom_stats;

/* Replaces `python3 reordering.py --m M --v V [--maxiter]` (reordering.py:
 * 183-265) and the reload of temp.mtx that follows it (main.cu:268): the
 * VENOM row reorder, then the loader's re-sort and the %.2f/%.1e text
 * quantisation of the values. */
int vnm_venom_reorder(vnm_graph *g, int V, int maxiter, vnm_venom_stats *st);

/* The 16 numbers of the reference's stdout CSV line (main.cu:331-335). */
typedef struct {
    float   nrows, nnz, density;
    float   init_size, final_size, init_cnt, final_cnt, improve_rate;
    int32_t init_mb, final_mb, init_vnm, final_vnm;
    int32_t nm_rounds, vnm_rounds, total_rounds;
    float   milliseconds;
} vnm_hybrid_stats;

/* Replaces hybrid_reordering (main.cu:195-337) with the temp.mtx / output.txt
 * process round trips done in device memory.  init_mb/init_vnm evaluate the
 * input matrix (the reference reads a stale ./temp.mtx, main.cu:207). */
int vnm_hybrid(vnm_graph *g, int N, int V, int rounds, int sched, vnm_hybrid_stats *st);

/* Replaces find_best_2toX (main.cu:72-114) over schedules 0..nsched-1, each on
 * a fresh copy of the graph, the copies reordered together as one
 * vnm_nm_reorder_batch; improve-rate ties go to the lowest index (the
 * reference compares wall times). */
int vnm_find_best_sched(vnm_graph *g, int N, int nsched, int32_t *best);

/* Stable LSD radix sort of device arrays on the context stream: keys_out =
 * keys_in ordered by bits [0, end_bit) (equal keys keep their input order),
 * vals_out the values moved with them (both value pointers NULL: keys only).
 * No counterpart in the reference, whose VENOM row order is Python's stable
 * sorted() over per-row keys (reordering.py:219); this is the library's own
 * sort behind the VENOM, packing and proposer passes.  Outputs must not
 * alias inputs; n < 2^31. */
int vnm_sort_pairs(vnm_ctx *ctx, const uint64_t *keys_in, uint64_t *keys_out, const int32_t *vals_in,
                   int32_t *vals_out, int64_t n, int end_bit);

/* ---------------------------------------------------------------- V:N:M */

/* Entries kept by the V:N:M (N = 2) format: per V x M block, the 4-column
 * subset maximising sum_rows min(2, |row & S|) (ties: lexicographically
 * smallest subset), summed over blocks.  Not in the reference. */
int vnm_retained_nnz(vnm_graph *g, int V, int N, int64_t *out);

#define VNM_DT_F32   0
#define VNM_DT_BF16  1
#define VNM_DT_F16   2
#define VNM_DT_FP8   3   /* OCP e4m3, per-block FP32 scale */

typedef struct {
    int32_t V, M, nrows, ncols;
    int32_t nblocks;          /* stored V x M blocks                     */
    int32_t dtype;
    int64_t retained_nnz;
    int64_t bytes;            /* device bytes of the packed format, its
                                 SpMM value stream included              */
    int64_t spmm_bytes;       /* bytes of it one vnm_spmm reads: the value
                                 stream (row, B row and FP32 weight per
                                 kept value) and the work-item table     */
    int32_t items;            /* SpMM work items (whole rows of at most 64
                                 values, or 64-value chunks of longer
                                 rows)                                    */
    int32_t slots;            /* rows of partial sums: one per chunk of a
                                 row longer than 64 values, plus V
                                 per extra work unit of the sparse-MFMA
                                 part; a product with n columns needs
                                 slots * n floats of scratch             */
    int32_t dense_blocks;     /* blocks in the sparse-MFMA part          */
    int32_t dense_chunks;     /* their chunks of 16 blocks               */
} vnm_packed_info;

/* Replaces the cuSPARSELt prune/compress path (kernel/spmm.cu:402-447) and
 * the BSR layout of bsr_wmma (kernel/spmm.cu:44-101): packs the current
 * matrix into the VENOM format (per stored V x M block: 4 column selectors,
 * V x 2 values + 2-bit indices). */
int  vnm_pack(vnm_graph *g, int V, int dtype, vnm_packed **out);
/* vnm_pack with a 2:4 sparse-MFMA part, replacing cuSPARSELt's 2:4 matmul
 * (kernel/spmm.cu:402-469) where it pays: blocks keeping at least dense_min
 * values (0: none) are packed as v_smfmac_f32_16x16x64_bf16 operands, 16
 * blocks of one V-row block side by side as K = 64, and vnm_spmm multiplies
 * them on the sparse matrix cores before the row entries of the other
 * blocks.  BF16 values only, V a multiple of 16; the product then takes BF16
 * B, and non-finite B values reach every row of a dense block's 16-row slice
 * (0 x Inf), not only the rows that keep a value.  Matrices with such a part
 * are not accepted by vnm_batch_create. */
int  vnm_pack_ex(vnm_graph *g, int V, int dtype, int dense_min, vnm_packed **out);
int  vnm_packed_info_get(const vnm_packed *p, vnm_packed_info *info);
void vnm_packed_destroy(vnm_packed *p);

/* Replaces evalCuSPARSELtMatmul / bsr_wmma's product (kernel/spmm.cu:44-101,
 * 324-500): C[nrows x n] = A_vnm * B[ncols x n], B and C row-major device
 * arrays (ldb, ldc in elements), B in b_dtype (F32, BF16 or FP8 with one FP32
 * scale per 128 rows in b_scales), C FP32, FP32 accumulation. */
int vnm_spmm(vnm_packed *p, const void *B, int ldb, int n, int b_dtype,
             const float *b_scales, float *C, int ldc);

/* vnm_spmm without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Bit for bit the
 * result of vnm_spmm.  The partial sums of rows cut into work-item chunks go
 * to a static device array of 8 Mi floats (32 MiB) that each device holds
 * until the process ends; every call on a device shares it, so calls that
 * overlap on one GPU (other contexts or streams) race.  A product needing
 * more (vnm_packed_info.slots * n floats) returns VNM_E_ARG.  The launch
 * shape is the one vnm_spmm chose or vnm_spmm_set_shape fixed for
 * (b_dtype, n), else 8 bytes per lane. */
int vnm_spmm_noalloc(vnm_packed *p, const void *B, int ldb, int n, int b_dtype,
                     const float *b_scales, float *C, int ldc);

/* Launch shape of vnm_spmm for (b_dtype, n): each lane gathers
 * cols_per_lane consecutive values of a B row per entry (2 to 16 bytes) and
 * the columns split into tiles of 64 * cols_per_lane; when the tile count
 * divides 8, each tile runs on its own XCDs.  Not in the reference (its
 * kernels have fixed shapes).  By default the first vnm_spmm call of a shape
 * times every usable value and keeps the fastest; set_shape with a value in
 * {1, 2, 4, 8, 16} fixes it, 0 clears it; get_shape gives 0 when the shape
 * has not been chosen yet. */
int vnm_spmm_set_shape(vnm_packed *p, int b_dtype, int n, int cols_per_lane);
int vnm_spmm_get_shape(const vnm_packed *p, int b_dtype, int n, int32_t *cols_per_lane);

/* GCN aggregation layer (the consumer of gnn/single/pyg_gcn.py:41-74,
 * aggr='add'): Y = A_vnm (X W) + bias, X: ncols x K BF16 (leading dimension
 * ldx), W: K x N BF16 row-major, bias: N floats or NULL, Y: nrows x N FP32.
 * X W runs on MFMA (FP32 accumulation, BF16 result kept with p), then the
 * SpMM with the bias added in its row writes.  K % 16 == 0, N % 64 == 0. */
int vnm_gcn_layer(vnm_packed *p, const uint16_t *X, int ldx, int K, const uint16_t *W, int N,
                  const float *bias, float *Y, int ldy);

/* Batched SpMM: C_g = A_g B_g (+ bias_g) for `count` packed matrices of one
 * value dtype and uneven sizes, all with n output columns, in one persistent
 * launch captured with its row pre-fill as a HIP graph.  Not in the
 * reference.  cols_per_lane = 0 tunes the launch shape on the batch itself.
 * The operand pointers are fixed at creation; vnm_batch_run replays the graph
 * on the context stream.  b_scales and bias may be NULL (per-entry NULLs
 * allowed for bias). */
typedef struct vnm_batch vnm_batch;
int  vnm_batch_create(vnm_ctx *ctx, vnm_packed *const *packed, int count, const void *const *B,
                      const int32_t *ldb, int n, int b_dtype, const float *const *b_scales,
                      float *const *C, const int32_t *ldc, const float *const *bias,
                      int cols_per_lane, vnm_batch **out);
int  vnm_batch_run(vnm_batch *b);
int  vnm_batch_info(const vnm_batch *b, int32_t *cols_per_lane, int32_t *units, int32_t *zero_rows);
void vnm_batch_destroy(vnm_batch *b);

/* ---------------------------------------------------------------- learned proposer */

/* A graph network that ranks candidate column pairs for the N:M reorder so
 * that only the top ones are scored exactly.  An addition of this library:
 * the reference has no learned component; the exact gain it is trained on
 * and judged by is GetSegSwapScore's (reorder/reorder.h:75-131).
 * Network: 16 features per node, three neighbour-mean layers of width 64
 * (BF16 activations and weights, FP32 accumulation), a 34-output head; pair
 * score s(u, v) = Y0[u] + Y1[v] + sum_j Y[2+j][u] * Y[18+j][v].  22784 FP32
 * parameters (layout: W0 [32][64], b0, W1 [128][64], b1, W2 [128][64], b2,
 * Wh [64][64], bh). */
#define VNM_LP_NPARAMS 22784
typedef struct vnm_lp_model vnm_lp_model;

/* params == NULL: seeded He-normal initialisation. */
int  vnm_lp_model_create(vnm_ctx *ctx, const float *params, uint64_t seed, vnm_lp_model **out);
void vnm_lp_model_destroy(vnm_lp_model *m);
int  vnm_lp_model_get(const vnm_lp_model *m, float *params);

/* Input features of the current matrix (nrows x 16, FP32 host array). */
int vnm_lp_features(vnm_graph *g, int N, float *X);
/* Head outputs (nrows x 64, FP32 host array); bf16 = 0 runs the FP32 path. */
int vnm_lp_forward(vnm_lp_model *m, vnm_graph *g, int N, int bf16, float *head);
/* Mean squared error of s(u_i, v_i) against target[i] (pairs = 2*count
 * current ids) and, when grad != NULL, its gradient (VNM_LP_NPARAMS). */
int vnm_lp_loss_grad(vnm_lp_model *m, vnm_graph *g, int N, int bf16, const int32_t *pairs,
                     const float *target, int count, float *loss, float *grad);
/* One Adam step (BF16 forward, FP32 backward) on one batch of pairs. */
int vnm_lp_train_step(vnm_lp_model *m, vnm_graph *g, int N, const int32_t *pairs,
                      const float *target, int count, float lr, float *loss);

/* Exact gain matrices of segment pairs (segp[k], segq[k]), k < count, from
 * the entries of the two segments; same values and layout as vnm_swap_score
 * (2*M*M ints per pair).  segp[k] != segq[k]. */
int vnm_pair_score(vnm_graph *g, int N, const int32_t *segp, const int32_t *segq, int count,
                   int32_t *scores);

#define VNM_LP_LEARNED    0  /* network proposals; exhaustive window when none pays */
#define VNM_LP_RANDOM     1  /* seeded random proposals (ablation)                 */
#define VNM_LP_EXHAUSTIVE 2  /* every window pair scored exactly, count heuristic  */

typedef struct {
    int32_t init_cnt, final_cnt;     /* rows with more than N ones, summed   */
    int32_t init_size, final_size;   /* segments holding such rows           */
    int32_t iters, swaps, fallbacks, gain;
    int64_t exact_pairs;             /* segment pairs scored exactly         */
    float   ms;
} vnm_lp_stats;

/* Learned N:M reorder: per iteration the top `topk` donor segments each
 * propose one pair from a window of `window` acceptor segments; the pairs
 * are scored exactly and applied greedily by exact gain, segment-disjoint,
 * positive gains only, so final_cnt <= init_cnt.  Relabels the graph like
 * vnm_nm_reorder.  m may be NULL for modes 1 and 2. */
int vnm_lp_reorder(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window,
                   int maxiter, uint64_t seed, vnm_lp_stats *st);

/* vnm_lp_reorder of count graphs of one context as one launch chain: one
 * iteration of every graph is captured as a HIP graph of 8 parallel branches
 * (every kernel reads its sizes and flags from its graph's device-side
 * control block) and replayed until every run is over or maxiter iterations
 * have run, with one host read of the runs' done flags per replay (and a new
 * capture of the running graphs once at most half of the captured ones
 * run).  st[i] and graph i's relabelling are those of vnm_lp_reorder on
 * graph i alone with the same seed.  No counterpart in the reference; for
 * batches of graphs such as the inputs of one GCN training set. */
int vnm_lp_reorder_batch(vnm_graph *const *graphs, int count, const vnm_lp_model *m, int N, int mode,
                         int topk, int window, int maxiter, uint64_t seed, vnm_lp_stats *st);

/* hybrid_reordering (main.cu:195-337) with vnm_lp_reorder (mode, topk,
 * window, maxiter) as its N:M step; VENOM steps and stats as vnm_hybrid
 * (nm_rounds counts proposer iterations). */
int vnm_hybrid_learned(vnm_graph *g, const vnm_lp_model *m, int N, int V, int rounds, int mode,
                       int topk, int window, int maxiter, vnm_hybrid_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* VNM_H */
