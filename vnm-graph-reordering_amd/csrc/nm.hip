// N:M reorder engine for gfx950 (wave64).
//
// Replaces the reference's per-call kernels and host loop
// (reorder/reorder.h:23-317 kernels, :392-797 helpers, :963-1100 and
// :1805-2474 schedules, mtx/Mtx.h:437-478 per-round rebuild).
//
// Layout.  At the start of a run the nonzero entries are bucketed into a CSC
// over the run-start ids (cptr/crow).  The run keeps the relabel as a pair of
// device arrays perm (run-start id -> current id) and inv (current id ->
// run-start id) instead of rewriting the COO every round: the bitmap of
// (row, segment s) is the OR over b < M of bit (M-1-b) for every entry of
// run-start column inv[s*M + b] in that row.  Rows are never renamed inside
// a run: a symmetric relabel only permutes rows, which changes no count or
// gain, so rows are identified by run-start id.
//
// Per-segment "tables": a slot holds one byte per row (R bytes, 4 rows per
// 32-bit word), filled by atomicOr from the segment's M column lists, read
// back once per entry ("owner" = the entry carrying the row's leading bit,
// so each row is counted once), then cleared by the same entries.
//
// Kernels:
//   k_col_hist / k_scan_* / k_col_scatter : CSC build (counting sort)
//   k_locate_all  : cnt/nzr for every segment (LocateinvalidSegment /
//                   LocateAllZeroSegment, reorder/reorder.h:262-317)
//   k_score_rows  : batched gain matrix of one segment against many, from the
//                   slot bytes with wave ballots and scalar popcounts
//                   (GetSegSwapScore, reorder/reorder.h:75-131)
//   k_pq_pop_order: libstdc++ priority_queue pop order (reorder/swappair.h)
//   k_gr_*        : one greedy round as a chain of kernels (see the greedy
//                   section below)
//   k_relabel     : the Reorder kernel (reorder/reorder.h:23-35), applied once
#include "graph.h"
#include "sortlib.h"
#include "hip_common.h"
#include "tables.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace vnm {

// ------------------------------------------------------------ small kernels

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_iota_i32(int32_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (int32_t)i;
}

__global__ void k_col_hist(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                           const float *__restrict__ val, int64_t nnz, int32_t R,
                           int32_t *__restrict__ cnt1) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        int c = col[i], r = row[i];
        if (val[i] != 0.0f && c < R && r < R) atomicAdd(&cnt1[c + 1], 1);
    }
}

__global__ void k_col_scatter(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                              const float *__restrict__ val, int64_t nnz, int32_t R,
                              int32_t *__restrict__ cursor, int32_t *__restrict__ crow) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        int c = col[i], r = row[i];
        if (val[i] != 0.0f && c < R && r < R) crow[atomicAdd(&cursor[c], 1)] = r;
    }
}

// Inclusive scan of n ints in place, three phases (1024-element tiles).
constexpr int SCAN_TILE = 1024;

__device__ int block_incl_scan(int v, int *sh) {
    // 1024 threads = 16 waves; wave scan then scan of wave totals
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    if (lane == 63) sh[w] = v;
    __syncthreads();
    if (w == 0) {
        int x = lane < 16 ? sh[lane] : 0;
        for (int d = 1; d < 16; d <<= 1) {
            int t = __shfl_up(x, d, 64);
            if (lane >= d) x += t;
        }
        if (lane < 16) sh[lane] = x;
    }
    __syncthreads();
    if (w > 0) v += sh[w - 1];
    __syncthreads();
    return v;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(int32_t *p, int64_t n, int32_t *tile_tot) {
    __shared__ int sh[16];
    int64_t i = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    int v = i < n ? p[i] : 0;
    v = block_incl_scan(v, sh);
    if (i < n) p[i] = v;
    if (threadIdx.x == SCAN_TILE - 1) tile_tot[blockIdx.x] = v;
}

__global__ __launch_bounds__(1024) void k_scan_totals(int32_t *tot, int ntiles) {
    __shared__ int sh[16];
    int carry = 0;
    for (int base = 0; base < ntiles; base += SCAN_TILE) {
        int i = base + threadIdx.x;
        int v = i < ntiles ? tot[i] : 0;
        v = block_incl_scan(v, sh);
        if (i < ntiles) tot[i] = v + carry;
        __shared__ int last;
        if (threadIdx.x == SCAN_TILE - 1) last = v;
        __syncthreads();
        carry += last;
        __syncthreads();
    }
}

__global__ void k_scan_add(int32_t *p, int64_t n, const int32_t *tile_tot) {
    int64_t i = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    if (blockIdx.x > 0 && i < n) p[i] += tile_tot[blockIdx.x - 1];
}

// ------------------------------------------------------------ locate (all segments)

constexpr int LOC_T = 256;

__global__ __launch_bounds__(LOC_T) void k_locate_all(
    int R, int M, int N, int nseg, const int32_t *__restrict__ inv,
    const int32_t *__restrict__ cptr, const int32_t *__restrict__ crow, uint32_t *tab,
    int32_t *__restrict__ cnt, int32_t *__restrict__ nzr, int32_t *__restrict__ queue) {
    __shared__ ColSpan cols[16];
    __shared__ int pre[17];
    __shared__ int sel[1], cur;
    __shared__ int c_inv, c_nz;
    __shared__ RowHash H;
    uint32_t *slot = tab + (size_t)blockIdx.x * (R >> 2);
    hash_reset(H);
    // segments from a work counter (*queue zero at launch): sizes are
    // heavy-tailed and cluster by index
    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(queue, 1);
        __syncthreads();
        const int s = cur;
        if (s >= nseg) break;
        if (threadIdx.x == 0) { sel[0] = s; c_inv = 0; c_nz = 0; }
        __syncthreads();
        int total = table_columns(sel, 1, M, inv, cptr, cols, pre);
        // row bitmaps: the LDS hash for small segments, the global slot else
        const bool small = total <= HASH_MAX;
        if (small) hash_scatter(H, M, cols, pre, M, total, crow);
        else table_scatter(slot, M, R, cols, pre, M, total, crow);
        __syncthreads();
        int my_inv = 0, my_nz = 0;
        // TAB_U entries per thread resolved before their loads are issued
        for (int e0 = threadIdx.x; e0 < total; e0 += blockDim.x * TAB_U) {
            int t[TAB_U], r[TAB_U];
            uint32_t bits[TAB_U];
            table_entries(cols, pre, M, total, crow, e0, blockDim.x, t, r);
#pragma unroll
            for (int u = 0; u < TAB_U; ++u)
                bits[u] = t[u] < 0 ? 0u : small ? (hash_get(H, (uint32_t)r[u]) & 0xFFu) : slot_byte(slot, 0, R, r[u]);
#pragma unroll
            for (int u = 0; u < TAB_U; ++u)
                if (t[u] >= 0 && (bits[u] >> (M - 1 - t[u])) == 1u) {  // leading bit: the row's owner
                    my_nz += 1;
                    my_inv += popc32(bits[u]) > N;
                }
        }
        if (my_inv) atomicAdd(&c_inv, my_inv);
        if (my_nz) atomicAdd(&c_nz, my_nz);
        __syncthreads();
        if (small) hash_reset(H);
        else table_clear(slot, M, R, cols, pre, M, total, crow);
        if (threadIdx.x == 0) { cnt[s] = c_inv; nzr[s] = c_nz; }
        __syncthreads();
    }
}

// ------------------------------------------------------------ batched scorer

// Fill (or clear) the slots of `nsegs` segments: one workgroup per segment.
__global__ __launch_bounds__(256) void k_fill_slots(int R, int M, const int32_t *__restrict__ segs,
                                                     const int32_t *__restrict__ inv,
                                                     const int32_t *__restrict__ cptr,
                                                     const int32_t *__restrict__ crow,
                                                     uint32_t *tab, int clear) {
    __shared__ ColSpan cols[16];
    __shared__ int pre[17];
    __shared__ int sel[1];
    if (threadIdx.x == 0) sel[0] = segs[blockIdx.x];
    __syncthreads();
    int total = table_columns(sel, 1, M, inv, cptr, cols, pre);
    uint32_t *slot = tab + (size_t)blockIdx.x * (R >> 2);
    if (clear) table_clear(slot, M, R, cols, pre, M, total, crow);
    else table_scatter(slot, M, R, cols, pre, M, total, crow);
}

// Gain matrix of slot 0 against slots 1..ntgt over all R rows.
// Each wave takes 64 consecutive rows: lane i holds row i's bytes l (slot 0)
// and r (target slot); ballots give the bit planes L_b, R_b and the classes
// A = [popc l == N+1], B = [popc l == N], C = [popc r == N], D = [popc r ==
// N+1]; lane (b1, b2) then accumulates the closed form
//   S_l += popc(L_b1 & ~R_b2 & A) - popc(~L_b1 & R_b2 & B)
//   S_r += popc(~L_b1 & R_b2 & D) - popc(L_b1 & ~R_b2 & C)
// which equals the per-row sum of GetSegSwapScore's lgain/rgain
// (reorder/reorder.h:108-122; rows with l = r = 0 add nothing for N >= 1).
constexpr int SC_WAVES = 4;
constexpr int SC_ROWS = SC_WAVES * 64;

__global__ __launch_bounds__(SC_ROWS) void k_score_rows(int R, int M, int N, int ntgt,
                                                         const uint32_t *__restrict__ tab,
                                                         int32_t *__restrict__ score) {
    __shared__ int part[SC_WAVES][2][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int MM = M * M;
    const int b1 = lane / M, b2 = lane % M;
    const bool act = lane < MM;
    const size_t words = (size_t)(R >> 2);
    for (int k = blockIdx.y; k < ntgt; k += gridDim.y) {
        int acc_l = 0, acc_r = 0;
        for (int base = blockIdx.x * SC_ROWS; base < R; base += gridDim.x * SC_ROWS) {
            int r = base + w * 64 + lane;
            uint32_t l = 0, rr = 0;
            if (r < R) {
                l = (tab[r >> 2] >> ((r & 3) * 8)) & 0xFFu;
                rr = (tab[(size_t)(k + 1) * words + (r >> 2)] >> ((r & 3) * 8)) & 0xFFu;
            }
            const int pl = popc32(l), pr = popc32(rr);
            const uint64_t A = __ballot(pl == N + 1), B = __ballot(pl == N);
            const uint64_t C = __ballot(pr == N), D = __ballot(pr == N + 1);
            if ((A | B | C | D) == 0) continue;  // no row here can change its class
            uint64_t Lb = 0, Rb = 0;
            for (int b = 0; b < M; ++b) {
                uint64_t lm = __ballot((l >> (M - 1 - b)) & 1u);
                uint64_t rm = __ballot((rr >> (M - 1 - b)) & 1u);
                if (b == b1) Lb = lm;
                if (b == b2) Rb = rm;
            }
            if (act) {
                acc_l += __popcll(Lb & ~Rb & A) - __popcll(~Lb & Rb & B);
                acc_r += __popcll(~Lb & Rb & D) - __popcll(Lb & ~Rb & C);
            }
        }
        part[w][0][lane] = acc_l;
        part[w][1][lane] = acc_r;
        __syncthreads();
        if (w == 0 && act) {
            int sl = 0, sr = 0;
            for (int q = 0; q < SC_WAVES; ++q) { sl += part[q][0][lane]; sr += part[q][1][lane]; }
            if (sl) atomicAdd(&score[(size_t)k * 2 * MM + lane * 2 + 0], sl);
            if (sr) atomicAdd(&score[(size_t)k * 2 * MM + lane * 2 + 1], sr);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ heap pop order

// One thread per list: dense ranks of the list's M*M keys, then the same
// heap code as the greedy pushes every pair and pops them all.
__global__ void k_pq_pop_order(const int32_t *__restrict__ scores, int count, int M,
                               int32_t *__restrict__ order) {
    extern __shared__ uint32_t hsm[];  // thread t's heap at [t * HEAP_WORDS, (t + 1) * HEAP_WORDS)
    const int MM = M * M;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int32_t *sc = scores + (size_t)q * 2 * MM;
    HeapRef H{reinterpret_cast<uint16_t *>(hsm + threadIdx.x * HEAP_WORDS), 0};
    for (int i = 0; i < MM; ++i) {
        const int64_t ki = swap_key(sc[2 * i], sc[2 * i + 1]);
        int r = 0;
        for (int j = 0; j < MM; ++j) r += swap_key(sc[2 * j], sc[2 * j + 1]) < ki;
        heap_push(H, (uint32_t)((r << 8) | i));
    }
    for (int k = 0; k < MM; ++k) order[(size_t)q * MM + k] = (int32_t)(heap_pop(H) & 0xFFu);
}

// ------------------------------------------------------------ relabel

__global__ void k_relabel(int32_t *__restrict__ row, int32_t *__restrict__ col, int64_t nnz,
                          const int32_t *__restrict__ p, int32_t R) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        int r = row[i], c = col[i];
        if (r < R) row[i] = p[r];
        if (c < R) col[i] = p[c];
    }
}

// Largest row and column id of the entries: one atomic per wave.
__global__ __launch_bounds__(256) void k_max_ids(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                                 int64_t nnz, int32_t *__restrict__ out) {
    int mr = -1, mc = -1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        mr = max(mr, row[i]);
        mc = max(mc, col[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mr = max(mr, __shfl_xor(mr, o));
        mc = max(mc, __shfl_xor(mc, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], mr);
        atomicMax(&out[1], mc);
    }
}

// total[i] = p[total[i]] over the ids at creation; ids at or past the p's
// length (empty rows a reload's smaller shape cut off) stay as they are.
__global__ void k_compose(int32_t *__restrict__ total, int64_t n, const int32_t *__restrict__ p, int32_t R) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t t = total[i];
        if (t < R) total[i] = p[t];
    }
}

// ------------------------------------------------------------ greedy
//
// One N:M run is a chain of kernels per round on the context's stream:
//
//   k_gr_select   (1 workgroup)   commit the last round's recount, stop or
//                                 choose the round's segments, their columns
//   k_gr_fill_a   (grid)          clear records B (last recount), fill A
//   k_gr_score    (grid)          gains of every pair from records A
//   k_gr_greedy   (1 workgroup)   priority queues, sequential greedy, commit
//                                 to perm/inv, columns after the commit
//   k_gr_fill_b   (grid)          clear records A, fill B
//   k_gr_count    (grid)          cnt / nzr of the round's segments from B
//
// The passes over the round's entries, which hub segments make large, run
// on every CU; the sequential parts keep one workgroup.  Kernel boundaries
// order everything, so no workgroup waits on another.  The host enqueues
// rounds in batches; once k_gr_select sets `done` every later kernel
// returns at entry.  Records: 16 bytes per row, byte j = the row's M-bit
// bitmap in slot j of the round (j < KMAX); A holds the segments' columns
// before the round's swaps, B after; both are zero between runs.

constexpr int GT = 1024;         // threads of the single-workgroup kernels
constexpr int NW = GT / 64;      // their waves
constexpr int PT = 256;          // threads of the pass kernels
constexpr int PNW = PT / 64;
constexpr int KMAX = 9;          // primary + up to M targets (M <= 8)
constexpr int MMAX = 8;
constexpr int NPAIR_T = KMAX * (KMAX - 1) / 2 + 1;  // unordered pairs + one self pair
constexpr int NQ = (KMAX - 1) * (KMAX - 1);         // ordered pairs the greedy can visit
constexpr int POPP = 16;                            // fewest pops precomputed per ordered pair
static_assert(NQ <= 64, "one heap thread per ordered pair, all in wave 0");
constexpr int NACC = KMAX * MMAX + NPAIR_T * 2 * MMAX * MMAX;  // E and both, flattened

// The columns of one set of records: slot j, bit b = column j*M + b.
struct ColSet {
    ColSpan cols[KMAX * MMAX];
    int pre[KMAX * MMAX + 1];
    int cb[KMAX + 1];  // 64-entry chunks of slot j: [cb[j], cb[j+1])
    int K, total;
};

// Device state of one run.
struct GreedyState {
    // parameters
    int R, M, N, nseg, sched, round_cap;
    int64_t trace_cap;
    int32_t *trace;  // optional: 9 ints per swap
    // run state
    int round, L, totcnt, maxiter, init_size, init_cnt, status, done, pending;
    int npre;  // pops to precompute per ordered pair this round (from the last round's need)
    int K, self_pair, size1;
    int sel[KMAX], selcnt[KMAX];
    int64_t trace_len, work;
    ColSet S, C;  // columns of records A (score) and B (recount)
    int acc[NACC];
    int newcnt[KMAX], newnzr[KMAX];
    uint64_t clk[16];
};

// Arguments of the round kernels for one graph: pointers fixed for the
// graph.  The kernels take graph 0's by value and an array of all of them in
// device memory, and serve graph blockIdx.y, so one launch chain runs the
// greedy of many graphs at once (a single graph reads no argument memory).
struct NMArgs {
    GreedyState *st;
    const int32_t *cptr, *crow;
    int32_t *perm, *inv, *cnt, *nzr;
    uint32_t *rta, *rtb;
};

// LDS of the single-workgroup kernels.
struct GreedyShared {
    int sel[KMAX];
    int selcnt[KMAX];
    int K, self_pair, size1, cont, status, L;
    ColSet cs;
    int E[KMAX][MMAX];
    // per unordered pair (j < j2) and one self pair: the "both rows" sums
    //   Bl[b1*M+b2] = sum over rows in both of y_b2 * (x_b1 ? A : B)
    //   Br[b1*M+b2] = sum over rows in both of x_b1 * (y_b2 ? D : C)
    int both[NPAIR_T][2][MMAX * MMAX];
    // per ordered pair q: the heap, its first
    // npre pops, the next of those not yet consumed, and the heap's size
    uint32_t heap[NQ][HEAP_WORDS];
    uint8_t popped[MMAX * MMAX][NQ];
    int qcur[NQ];
    int heapn[NQ];
    int npre, heap_pops;  // pops precomputed this round; pops the greedy took from heaps
    int maxk;             // precomputed pops the greedy used, most over the ordered pairs
    int swv1[64], swv2[64], swlg[64], swrg[64], swp[64], swt[64], swpc[64], swtc[64];
    int nsw;
    uint64_t wbest[NW];
    uint64_t topk[KMAX + 1];
    int selthr[KMAX + 1], nselthr;  // threads whose segments hold the round's top keys
    int64_t wkeys[NW][64];
};

__device__ __forceinline__ uint64_t key_fwd(int s, int c) {
    return c > 0 ? ((uint64_t)(uint32_t)c << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)s) : 0ull;
}
__device__ __forceinline__ uint64_t key_rev(int s, int c) {
    return c > 0 ? ((uint64_t)(uint32_t)(0x7FFFFFFF - c) << 32) | (uint64_t)(uint32_t)s : 0ull;
}
__device__ __forceinline__ int seg_of_key(uint64_t k, bool fwd) {
    uint32_t lo = (uint32_t)k;
    return fwd ? (int)(0xFFFFFFFFu - lo) : (int)lo;
}
__device__ __forceinline__ int cnt_of_key(uint64_t k, bool fwd) {
    const int hi = (int)(uint32_t)(k >> 32);
    return fwd ? hi : 0x7FFFFFFF - hi;
}

// Maximum over the wave, every lane gets it: a running maximum along DPP
// row shifts (1, 2, 4, 8 within rows of 16), then row broadcasts 15 and 31;
// lane 63 ends with the maximum.  Lanes without a DPP source read 0, the
// identity of an unsigned maximum.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_max_step(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWS, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROWS, 0xF, false);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    return o > v ? o : v;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    v = dpp_max_step<0x111, 0xF>(v);  // row_shr:1
    v = dpp_max_step<0x112, 0xF>(v);  // row_shr:2
    v = dpp_max_step<0x114, 0xF>(v);  // row_shr:4
    v = dpp_max_step<0x118, 0xF>(v);  // row_shr:8
    v = dpp_max_step<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
    v = dpp_max_step<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

// Top k (k <= KMAX + 1) of n keys in LDS (0 = none) into S.topk, by wave 0
// alone: each lane keeps a descending list of its keys (j = lane, lane + 64,
// ...), then k wave maxima merge the lists.  The other waves skip it (one
// wave's short chains instead of sixteen waves sharing four SIMDs).
__device__ __forceinline__ void wave0_topk(const uint64_t *keys, int n, int k, GreedyShared &S) {
    const int lane = threadIdx.x & 63;
    if (threadIdx.x >= 64) return;
    uint64_t lst[KMAX + 1];
#pragma unroll
    for (int i = 0; i <= KMAX; ++i) lst[i] = 0;
    for (int j = lane; j < n; j += 64) {
        uint64_t x = keys[j];
        if (x <= lst[KMAX]) continue;
#pragma unroll
        for (int i = 0; i <= KMAX; ++i) {
            const uint64_t hi = lst[i] > x ? lst[i] : x;
            x = lst[i] > x ? x : lst[i];
            lst[i] = hi;
        }
    }
    for (int it = 0; it < k; ++it) {
        const uint64_t m = wave_max_u64(lst[0]);
        if (m != 0 && lst[0] == m) {
#pragma unroll
            for (int i = 0; i < KMAX; ++i) lst[i] = lst[i + 1];
            lst[KMAX] = 0;
        }
        if (lane == 0) S.topk[it] = m;
    }
}

__device__ __forceinline__ uint64_t block_max_u64(uint64_t v, GreedyShared &S) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_max_u64(v);
    if (lane == 0) S.wbest[w] = v;
    __syncthreads();
    if (w == 0) {
        uint64_t x = lane < NW ? S.wbest[lane] : 0ull;
        x = wave_max_u64(x);
        if (lane == 0) S.wbest[0] = x;
    }
    __syncthreads();
    uint64_t r = S.wbest[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ int tri_index(int a, int b, int K) {
    // a < b < K
    return a * (2 * K - a - 1) / 2 + (b - a - 1);
}

// lgain/rgain of element e = b1*M + b2 of ordered pair (a, b) of slots:
//   score(a, b)[b1][b2] = (E_a[b1] - Bl(a,b)[b1][b2], E_b[b2] - Br(a,b)[b1][b2])
// and score(b, a) is score(a, b) transposed with the two gains exchanged.
__device__ __forceinline__ void pair_gain(const GreedyShared &S, int K, int M, int a, int b, int e,
                                          int &lg, int &rg) {
    const int b1 = e / M, b2 = e % M;
    if (a == b) {
        lg = S.E[a][b1] - S.both[NPAIR_T - 1][0][e];
        rg = S.E[a][b2] - S.both[NPAIR_T - 1][1][e];
    } else if (a < b) {
        const int pi = tri_index(a, b, K);
        lg = S.E[a][b1] - S.both[pi][0][e];
        rg = S.E[b][b2] - S.both[pi][1][e];
    } else {
        const int pi = tri_index(b, a, K);
        const int t = b2 * M + b1;
        lg = S.E[a][b1] - S.both[pi][1][t];
        rg = S.E[b][b2] - S.both[pi][0][t];
    }
}

// Ordered pair index of (a, b): (0, t) -> t-1; (i, j) with i, j >= 1 -> after.
__device__ __forceinline__ int q_index(int a, int b, int K) {
    if (a == 0) return b - 1;
    const int n = K - 1;
    return n + (a - 1) * (n - 1) + (b - 1 - (b > a ? 1 : 0));
}

__device__ __forceinline__ void q_pair(int q, int K, int &a, int &b) {
    const int n = K - 1;
    if (q < n) { a = 0; b = q + 1; return; }
    q -= n;
    a = 1 + q / (n - 1);
    int r = q % (n - 1);
    b = 1 + r + (r + 1 >= a ? 1 : 0);
}

// ---- column sets and record passes

// Chunk table of a column set (thread 0, then a barrier).
__device__ __forceinline__ void chunk_table(ColSet &cs, int K, int M) {
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int j = 0; j < K; ++j) {
            cs.cb[j] = acc;
            acc += (cs.pre[j * M + M] - cs.pre[j * M] + 63) / 64;
        }
        cs.cb[K] = acc;
        cs.K = K;
    }
    __syncthreads();
}

__device__ __forceinline__ int chunk_slot(const ColSet &cs, int K, int c) {
    int j = 0;
    while (j + 1 < K && cs.cb[j + 1] <= c) ++j;
    return j;
}

// Entry e of slot j: its column bit b and row r.
__device__ __forceinline__ void slot_entry(const ColSet &cs, const int32_t *crow, int j, int M, int e,
                                           int &b, int &r) {
    b = entry_col(cs.pre + j * M, M, e);
    const int t = j * M + b;
    r = crow[cs.cols[t].start + (e - cs.pre[t])];
}

__device__ __forceinline__ void copy_words(void *dst, const void *src, int bytes) {
    const int n = bytes / 4;
    for (int i = threadIdx.x; i < n; i += blockDim.x) ((int *)dst)[i] = ((const int *)src)[i];
}

__device__ __forceinline__ uint32_t rec_byte(const uint4 &v, int j) {
    const uint32_t w = (j >> 2) == 0 ? v.x : (j >> 2) == 1 ? v.y : v.z;
    return (w >> ((j & 3) * 8)) & 0xFFu;
}

// Set (CLEAR = false) or zero the records of this workgroup's share of the
// entries of a column set.
template <bool CLEAR>
__device__ __forceinline__ void rec_pass(uint32_t *rt, const ColSet &cs, int M, const int32_t *crow) {
    const int nc = cs.K * M, total = cs.total;
    const int stride = gridDim.x * blockDim.x;
    for (int e0 = blockIdx.x * blockDim.x + threadIdx.x; e0 < total; e0 += stride * TAB_U) {
        int t[TAB_U], r[TAB_U];
        table_entries(cs.cols, cs.pre, nc, total, crow, e0, stride, t, r);
#pragma unroll
        for (int u = 0; u < TAB_U; ++u) {
            if (t[u] < 0) continue;
            if constexpr (CLEAR) {
                *(uint4 *)(rt + 4 * (size_t)r[u]) = make_uint4(0u, 0u, 0u, 0u);
            } else {
                const int j = t[u] / M, b = t[u] % M;
                atomicOr(rt + 4 * (size_t)r[u] + (j >> 2), (1u << (M - 1 - b)) << ((j & 3) * 8));
            }
        }
    }
}

__device__ __forceinline__ void clock_add(GreedyState *st, int i, uint64_t t0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->clk[i] += __builtin_amdgcn_s_memrealtime() - t0;
}

// Records B of the last round's recount are cleared, records A of this
// round's segments filled.
__global__ __launch_bounds__(PT) void k_gr_fill_a(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ ColSet sa, sb;
    __shared__ int done;
    GreedyState *st = A.st;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    copy_words(&sa, &st->S, sizeof(ColSet));
    copy_words(&sb, &st->C, sizeof(ColSet));
    __syncthreads();
    const int M = st->M;
    rec_pass<true>(A.rtb, sb, M, A.crow);
    rec_pass<false>(A.rta, sa, M, A.crow);
    clock_add(st, 1, t0);
}

// Records A of this round are cleared, records B filled with the columns
// after the round's swaps.
__global__ __launch_bounds__(PT) void k_gr_fill_b(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ ColSet sa, sb;
    __shared__ int done;
    GreedyState *st = A.st;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    copy_words(&sa, &st->S, sizeof(ColSet));
    copy_words(&sb, &st->C, sizeof(ColSet));
    __syncthreads();
    const int M = st->M;
    rec_pass<true>(A.rta, sa, M, A.crow);
    rec_pass<false>(A.rtb, sb, M, A.crow);
    clock_add(st, 4, t0);
}

// After a run: both record sets back to zero (zeroing a clean record is
// harmless, so this needs no knowledge of where the run stopped).
__global__ __launch_bounds__(PT) void k_gr_cleanup(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ ColSet sa, sb;
    GreedyState *st = A.st;
    copy_words(&sa, &st->S, sizeof(ColSet));
    copy_words(&sb, &st->C, sizeof(ColSet));
    __syncthreads();
    const int M = st->M;
    rec_pass<true>(A.rta, sa, M, A.crow);
    rec_pass<true>(A.rtb, sb, M, A.crow);
}

// Gains of every pair of the round's segments from records A (the
// GetSegSwapScore sums, reorder/reorder.h:75-131, in closed form).  A wave
// takes a 64-entry chunk of one slot; the owner entries (the entry carrying
// the row's leading bit, one per row) are folded in with ballots, so each
// (b1, b2) counter takes one LDS atomic per chunk; each workgroup adds its
// nonzero sums to the device accumulators.
__global__ __launch_bounds__(PT) void k_gr_score(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ ColSet sa;
    __shared__ int E[KMAX][MMAX];
    __shared__ int both[NPAIR_T][2][MMAX * MMAX];
    __shared__ int done;
    GreedyState *st = A.st;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    copy_words(&sa, &st->S, sizeof(ColSet));
    for (int t = threadIdx.x; t < KMAX * MMAX; t += PT) (&E[0][0])[t] = 0;
    for (int t = threadIdx.x; t < NPAIR_T * 2 * MMAX * MMAX; t += PT) (&both[0][0][0])[t] = 0;
    __syncthreads();
    const int M = st->M, N = st->N, MM = M * M;
    const int K = sa.K, self_pair = st->self_pair;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nch = sa.cb[K];
    const int b1 = lane / M, b2 = lane % M;
    for (int c = blockIdx.x * PNW + w; c < nch; c += gridDim.x * PNW) {
        const int j = chunk_slot(sa, K, c);
        const int e1 = sa.pre[j * M + M];
        const int e = sa.pre[j * M] + (c - sa.cb[j]) * 64 + lane;
        uint32_t l = 0;
        uint4 rec = make_uint4(0u, 0u, 0u, 0u);
        bool act = false;
        if (e < e1) {
            int b, r;
            slot_entry(sa, A.crow, j, M, e, b, r);
            rec = *(const uint4 *)(A.rta + 4 * (size_t)r);
            l = rec_byte(rec, j);
            act = (l >> (M - 1 - b)) == 1u;  // the row's leading-bit entry
            if (!act) l = 0;
        }
        const int pl = popc32(l);
        const uint64_t Am = __ballot(act && pl == N + 1), Bm = __ballot(act && pl == N);
        uint64_t Lb1 = 0, Lb2 = 0, Lme = 0;
        for (int k = 0; k < M; ++k) {
            const uint64_t m = __ballot((l >> (M - 1 - k)) & 1u);
            if (k == b1) Lb1 = m;
            if (k == b2) Lb2 = m;
            if (k == lane) Lme = m;
        }
        if (lane < M) {
            const int v = __popcll(Lme & Am);
            if (v) atomicAdd(&E[j][lane], v);
        }
#pragma unroll
        for (int j2 = 0; j2 < KMAX; ++j2) {
            if (j2 <= j || j2 >= K) continue;
            const uint32_t rr = act ? rec_byte(rec, j2) : 0u;
            const int pr = popc32(rr);
            const uint64_t Cm = __ballot(pr == N), Dm = __ballot(pr == N + 1);
            uint64_t Rb = 0;
            for (int k = 0; k < M; ++k) {
                const uint64_t m = __ballot((rr >> (M - 1 - k)) & 1u);
                if (k == b2) Rb = m;
            }
            if (lane < MM) {
                const int vl = __popcll(Rb & ((Lb1 & Am) | (~Lb1 & Bm)));
                const int vr = __popcll(Lb1 & ((Rb & Dm) | (~Rb & Cm)));
                const int pi = tri_index(j, j2, K);
                if (vl) atomicAdd(&both[pi][0][lane], vl);
                if (vr) atomicAdd(&both[pi][1][lane], vr);
            }
        }
        if (self_pair && lane < MM) {  // seg1 == seg2: r = l for every row
            const int vl = __popcll(Lb2 & ((Lb1 & Am) | (~Lb1 & Bm)));
            const int vr = __popcll(Lb1 & ((Lb2 & Am) | (~Lb2 & Bm)));
            if (vl) atomicAdd(&both[NPAIR_T - 1][0][lane], vl);
            if (vr) atomicAdd(&both[NPAIR_T - 1][1][lane], vr);
        }
    }
    __syncthreads();
    if (blockIdx.x * PNW < nch) {
        for (int t = threadIdx.x; t < NACC; t += PT) {
            const int v = t < KMAX * MMAX ? (&E[0][0])[t] : (&both[0][0][0])[t - KMAX * MMAX];
            if (v) atomicAdd(&st->acc[t], v);
        }
    }
    clock_add(st, 2, t0);
}

// cnt / nzr of the round's segments after the swaps, from records B
// (locate_invalid_segment on the touched segments).
__global__ __launch_bounds__(PT) void k_gr_count(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ ColSet sb;
    __shared__ int cnt[KMAX], nzr[KMAX];
    __shared__ int done;
    GreedyState *st = A.st;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    copy_words(&sb, &st->C, sizeof(ColSet));
    if (threadIdx.x < KMAX) { cnt[threadIdx.x] = 0; nzr[threadIdx.x] = 0; }
    __syncthreads();
    const int M = st->M, N = st->N, K = sb.K;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nch = sb.cb[K];
    for (int c = blockIdx.x * PNW + w; c < nch; c += gridDim.x * PNW) {
        const int j = chunk_slot(sb, K, c);
        const int e1 = sb.pre[j * M + M];
        const int e = sb.pre[j * M] + (c - sb.cb[j]) * 64 + lane;
        bool act = false;
        uint32_t l = 0;
        if (e < e1) {
            int b, r;
            slot_entry(sb, A.crow, j, M, e, b, r);
            l = rec_byte(*(const uint4 *)(A.rtb + 4 * (size_t)r), j);
            act = (l >> (M - 1 - b)) == 1u;
        }
        const uint64_t o = __ballot(act), v = __ballot(act && popc32(l) > N);
        if (lane == 0) {
            if (o) atomicAdd(&nzr[j], (int)__popcll(o));
            if (v) atomicAdd(&cnt[j], (int)__popcll(v));
        }
    }
    __syncthreads();
    if (threadIdx.x < K) {
        if (cnt[threadIdx.x]) atomicAdd(&st->newcnt[threadIdx.x], cnt[threadIdx.x]);
        if (nzr[threadIdx.x]) atomicAdd(&st->newnzr[threadIdx.x], nzr[threadIdx.x]);
    }
    clock_add(st, 5, t0);
}

// Visited vertices of a round: bit a*M + v for vertex v of slot a (< 72).
struct Visited {
    uint64_t lo = 0, hi = 0;
    __device__ __forceinline__ bool has(int i) const {
        return ((i < 64 ? lo >> i : hi >> (i - 64)) & 1ull) != 0;
    }
    __device__ __forceinline__ void set(int i) {
        if (i < 64) lo |= 1ull << i;
        else hi |= 1ull << (i - 64);
    }
    // every vertex of slot a (M divides 64, so a slot's bits share a word)
    __device__ __forceinline__ bool full(int a, int M) const {
        const int i = a * M;
        const uint64_t m = (1ull << M) - 1;
        return (((i < 64 ? lo >> i : hi >> (i - 64))) & m) == m;
    }
};

// First element of ordered pair q (slots a -> b) in pop order whose two
// vertices are unvisited; -1 when the queue runs empty.  Skipped pops are
// consumed (a visited vertex stays visited for the round), so each queue
// resumes where its last search stopped: first the precomputed pops, one
// per lane and a ballot, then the heap itself.  Called by a whole wave with
// the same arguments in every lane.
__device__ __forceinline__ int first_free(GreedyShared &S, int q, int a, int b, int M, const Visited &vis) {
    const int lane = threadIdx.x & 63;
    const int npre = S.npre;  // <= 64
    const int k0 = S.qcur[q];
    if (k0 < npre) {
        const int k = k0 + lane;
        const int e = k < npre ? (int)S.popped[k][q] : 0;
        const bool ok = k < npre && !vis.has(a * M + e / M) && !vis.has(b * M + e % M);
        const uint64_t m = __ballot(ok);
        if (m) {
            const int i = __builtin_ctzll(m);
            S.qcur[q] = k0 + i;
            if (lane == 0) S.maxk = max(S.maxk, k0 + i + 1);  // precomputed pops this round used
            return __builtin_amdgcn_readlane(e, i);
        }
        S.qcur[q] = npre;
    }
    HeapRef H{reinterpret_cast<uint16_t *>(S.heap[q]), S.heapn[q]};
    while (H.n > 0) {
        const int e = (int)(heap_pop(H) & 0xFFu);
        S.heap_pops += 1;
        if (!vis.has(a * M + e / M) && !vis.has(b * M + e % M)) {
            S.heapn[q] = H.n;
            return e;
        }
    }
    S.heapn[q] = 0;
    return -1;
}

__device__ __forceinline__ bool list_before(int sa, int ca, int sb, int cb, bool rev) {
    // true when (sa, ca) is nearer the top of the invalid list than (sb, cb)
    if (!rev) return ca != cb ? ca > cb : sa < sb;
    return ca != cb ? ca < cb : sa > sb;
}

// The sequential part of one round (reorder/reorder.h:1837-1932), run by one
// wave in lockstep: every lane holds the same values and writes the same
// LDS words, and first_free spreads its search over the lanes.  Returns the
// number of swaps.
template <int M>  // 4 or 8: the per-element divisions by M become shifts
__device__ __forceinline__ int greedy_apply(GreedyShared &S, bool rev, bool drain) {
    const int K = S.K;
    int nsw = 0;
    Visited vis;
    auto add_swap = [&](int a, int b, int e, int lg, int rg, int pc, int tc) {
        const int n = nsw++;
        S.swv1[n] = S.sel[a] * M + e / M;
        S.swv2[n] = S.sel[b] * M + e % M;
        S.swlg[n] = lg; S.swrg[n] = rg;
        S.swp[n] = S.sel[a]; S.swt[n] = S.sel[b];
        S.swpc[n] = pc; S.swtc[n] = tc;
    };
    if (S.size1) {
        // one invalid segment: swap the top pair with the emptiest segment, no
        // visited filter (reorder/reorder.h:1837-1860)
        const int b = S.self_pair ? 0 : 1;
        const int e = S.popped[0][0];  // first pop of list 0
        int lg, rg;
        pair_gain(S, K, M, 0, b, e, lg, rg);
        add_swap(0, b, e, lg, rg, 0, 0);
        return nsw;
    }
    int cur = 0, pc = S.selcnt[0];
    int tgt[KMAX], tcnt[KMAX], nt = K - 1;
    for (int i = 0; i < nt; ++i) { tgt[i] = i + 1; tcnt[i] = S.selcnt[i + 1]; }
    while (true) {
        int upd[KMAX], ucnt[KMAX], nu = 0;
        for (int i = 0; i < nt; ++i) {
            if (pc <= 0 || vis.full(cur, M)) break;  // reorder/reorder.h:1878
            const int t = tgt[i];
            const int q = q_index(cur, t, K);
            int e = first_free(S, q, cur, t, M, vis);
            if (e < 0) {
                // top() of an empty queue: the reference would read the last
                // popped pair, which has a visited endpoint (:1908)
                S.status = VNM_E_UB_PATH;
                return nsw;
            }
            int lg, rg;
            pair_gain(S, K, M, cur, t, e, lg, rg);
            vis.set(cur * M + e / M);
            vis.set(t * M + e % M);
            pc -= lg;
            int tc = tcnt[i] - rg;
            add_swap(cur, t, e, lg, rg, pc, tc);
            if (tc > 0) { upd[nu] = t; ucnt[nu] = tc; ++nu; }
        }
        if (nu < 2) break;  // while (inv_seg_list.size() >= 2)
        for (int i = 1; i < nu; ++i) {
            int s = upd[i], c = ucnt[i], k = i - 1;
            while (k >= 0 && list_before(S.sel[s], c, S.sel[upd[k]], ucnt[k], rev)) {
                upd[k + 1] = upd[k]; ucnt[k + 1] = ucnt[k]; --k;
            }
            upd[k + 1] = s; ucnt[k + 1] = c;
        }
        cur = upd[0];
        pc = ucnt[0];
        nt = nu - 1;
        for (int i = 0; i < nt; ++i) {
            int src = drain ? (nu - 1 - i) : (i + 1);
            tgt[i] = upd[src];
            tcnt[i] = ucnt[src];
        }
    }
    return nsw;
}

// First kernel of a round: commit the last round's recount, decide whether
// the run goes on (reorder/reorder.h:1819-1826: rounds <= max(maxiter, list
// size), a list of at least one (two for schedules >= 8) segments), choose
// the round's segments and their columns.
__global__ __launch_bounds__(GT) void k_gr_select(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ GreedyShared S;
    GreedyState *st = A.st;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (tid == 0) S.cont = !st->done;
    __syncthreads();
    if (!S.cont) return;
    const int M = st->M, R = st->R, nseg = st->nseg, MM = M * M;
    const bool plain = st->sched >= 8;
    const bool rev = (st->sched == 2 || st->sched == 3);
    const bool drain = (st->sched == 1 || st->sched == 3);
    const bool tfwd = drain ? rev : !rev;  // traversal order uses the forward key
    const int min_size = plain ? 2 : 1;
    (void)lane; (void)w; (void)MM;
    // the last round's segments (distinct), one thread each
    if (tid == 0) { S.K = 0; S.L = 0; S.status = 0; }
    __syncthreads();
    const int pending = st->pending;
    if (pending && tid < st->K) {
        const int s = st->sel[tid];
        const int oc = ld_l2(&A.cnt[s]), nc = st->newcnt[tid];
        st_l2(&A.cnt[s], nc);
        st_l2(&A.nzr[s], st->newnzr[tid]);
        atomicAdd(&S.L, (nc != 0) - (oc != 0));
        atomicAdd(&S.status, nc - oc);
    }
    __syncthreads();
    if (tid == 0) {
        if (pending) {
            st->L += S.L;
            st->totcnt += S.status;
            st->round += 1;
            st->pending = 0;
        }
        S.cont = st->round <= st->maxiter && st->round <= st->round_cap && st->L >= min_size &&
                 st->status == 0;
        if (!S.cont) st->done = 1;
        S.L = st->L;
    }
    __syncthreads();
    if (!S.cont) return;
    const int Lr = S.L;
    const uint64_t tA = __builtin_amdgcn_s_memrealtime();

    // ---- phase A: the round's segments
    if (!plain && Lr == 1) {
        // the single invalid segment, and the segment with the most empty
        // rows, ties to the larger id (reorder/reorder.h:767-797)
        uint64_t best_l = 0, best_z = 0;
        for (int s = tid; s < nseg; s += GT) {
            int c = ld_l2(&A.cnt[s]);
            if (c != 0) best_l = 0x8000000000000000ull | (uint32_t)s;
            uint64_t z = ((uint64_t)(uint32_t)(R - ld_l2(&A.nzr[s])) << 32) | (uint32_t)s;
            best_z = z > best_z ? z : best_z;
        }
        uint64_t bl = block_max_u64(best_l, S);
        uint64_t bz = block_max_u64(best_z, S);
        if (tid == 0) {
            int top = (int)(uint32_t)bl, minseg = (int)(uint32_t)bz;
            S.size1 = 1;
            S.sel[0] = top;
            S.selcnt[0] = ld_l2(&A.cnt[top]);
            if (minseg == top) { S.K = 1; S.self_pair = 1; }
            else { S.K = 2; S.self_pair = 0; S.sel[1] = minseg; S.selcnt[1] = ld_l2(&A.cnt[minseg]); }
        }
        __syncthreads();
    } else {
        const int k = min(M, Lr - 1) + 1;  // targets + the primary itself
        // Top-k traversal keys in three steps.  Thread t scans the segments
        // s = t (mod GT) and keeps only its largest key; keys are distinct
        // (they hold s), so every key of the top k lies in a thread whose own
        // maximum is among the k largest thread maxima.  Those threads'
        // segments are then scored again, spread over the workgroup, and the
        // top k of them is the top k of all.
        uint64_t tk[KMAX + 1];
#pragma unroll
        for (int i = 0; i <= KMAX; ++i) tk[i] = 0;
        uint64_t best_l = 0;
        for (int s0 = tid; s0 < nseg; s0 += GT * 4) {
            int cv[4];  // four loads in flight per thread
#pragma unroll
            for (int u = 0; u < 4; ++u) cv[u] = s0 + u * GT < nseg ? ld_l2(&A.cnt[s0 + u * GT]) : 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = cv[u], s = s0 + u * GT;
                if (c == 0) continue;
                uint64_t lk = rev ? key_rev(s, c) : key_fwd(s, c);
                best_l = lk > best_l ? lk : best_l;
                const uint64_t x = tfwd ? key_fwd(s, c) : key_rev(s, c);
                tk[0] = x > tk[0] ? x : tk[0];
            }
        }
        uint64_t bl = block_max_u64(best_l, S);
        // the k largest thread maxima (one key per thread in LDS)
        uint64_t *keys = (uint64_t *)&S.wkeys[0][0];
        keys[tid] = tk[0];
        __syncthreads();
        wave0_topk(keys, GT, k, S);
        __syncthreads();
        if (tid == 0) {
            int n = 0;
            for (int i = 0; i < k; ++i)
                if (S.topk[i] != 0) S.selthr[n++] = seg_of_key(S.topk[i], tfwd) % GT;
            S.nselthr = n;
        }
        __syncthreads();
        // those threads' segments scored again by wave 0, then their top k
        const int per = (nseg + GT - 1) / GT;  // segments per thread, at most
        const int ncand = S.nselthr * per;
        if (tid < 64) {
            uint64_t lst[KMAX + 1];
#pragma unroll
            for (int i = 0; i <= KMAX; ++i) lst[i] = 0;
            for (int j = tid; j < ncand; j += 64) {
                const int s = S.selthr[j / per] + (j % per) * GT;
                const int c = s < nseg ? ld_l2(&A.cnt[s]) : 0;
                if (c == 0) continue;
                uint64_t x = tfwd ? key_fwd(s, c) : key_rev(s, c);
#pragma unroll
                for (int i = 0; i <= KMAX; ++i) {
                    const uint64_t hi = lst[i] > x ? lst[i] : x;
                    x = lst[i] > x ? x : lst[i];
                    lst[i] = hi;
                }
            }
            for (int it = 0; it < k; ++it) {
                const uint64_t m = wave_max_u64(lst[0]);
                if (m != 0 && lst[0] == m) {
#pragma unroll
                    for (int i = 0; i < KMAX; ++i) lst[i] = lst[i + 1];
                    lst[KMAX] = 0;
                }
                if (tid == 0) S.topk[it] = m;
            }
        }
        __syncthreads();
        if (tid == 0) {
            int primary = seg_of_key(bl, !rev);
            S.size1 = 0;
            S.self_pair = 0;
            S.sel[0] = primary;
            S.selcnt[0] = cnt_of_key(bl, !rev);
            int K = 1;
            for (int i = 0; i < k && K < k; ++i) {
                if (S.topk[i] == 0) break;
                int s = seg_of_key(S.topk[i], tfwd);
                if (s == primary) continue;
                S.sel[K] = s;
                S.selcnt[K] = cnt_of_key(S.topk[i], tfwd);
                ++K;
            }
            S.K = K;
        }
        __syncthreads();
    }


    const uint64_t tB = __builtin_amdgcn_s_memrealtime();
    const int K = S.K;
    const int total = table_columns(S.sel, K, M, A.inv, A.cptr, S.cs.cols, S.cs.pre);
    if (tid == 0) S.cs.total = total;
    chunk_table(S.cs, K, M);
    if (tid == 0) {
        st->clk[8] += tA - t0;
        st->clk[9] += tB - tA;
        st->clk[10] += __builtin_amdgcn_s_memrealtime() - tB;
    }
    copy_words(&st->S, &S.cs, sizeof(ColSet));
    if (tid < KMAX) {
        st->sel[tid] = S.sel[tid];
        st->selcnt[tid] = S.selcnt[tid];
    }
    if (tid == 0) {
        st->K = K;
        st->size1 = S.size1;
        st->self_pair = S.self_pair;
        st->work += total;
    }
    clock_add(st, 0, t0);
}

// The round's sequential part: priority queues of every ordered pair of the
// round's segments, the greedy (reorder/reorder.h:1837-1932), the commit to
// perm/inv, and the columns of the segments after the commit.
__global__ __launch_bounds__(GT) void k_gr_greedy(NMArgs a0, const NMArgs *__restrict__ args) {
    const NMArgs A = blockIdx.y == 0 ? a0 : args[blockIdx.y];
    __shared__ GreedyShared S;
    GreedyState *st = A.st;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    if (tid == 0) S.cont = !st->done;
    __syncthreads();
    if (!S.cont) return;
    const int M = st->M, R = st->R, nseg = st->nseg, MM = M * M;
    const bool plain = st->sched >= 8;
    const bool rev = (st->sched == 2 || st->sched == 3);
    const bool drain = (st->sched == 1 || st->sched == 3);
    const bool tfwd = drain ? rev : !rev;  // traversal order uses the forward key
    const int min_size = plain ? 2 : 1;
    (void)R; (void)nseg; (void)plain; (void)tfwd; (void)min_size;
    if (tid < KMAX) {
        S.sel[tid] = st->sel[tid];
        S.selcnt[tid] = st->selcnt[tid];
    }
    if (tid == 0) {
        S.K = st->K;
        S.size1 = st->size1;
        S.self_pair = st->self_pair;
        S.status = 0;
    }
    // the pair gains summed by k_gr_score; zeroed for the next round
    for (int t = tid; t < NACC; t += GT) {
        const int v = st->acc[t];
        st->acc[t] = 0;
        if (t < KMAX * MMAX) (&S.E[0][0])[t] = v;
        else (&S.both[0][0][0])[t - KMAX * MMAX] = v;
    }
    __syncthreads();

    const uint64_t tD = __builtin_amdgcn_s_memrealtime();
    // ---- phase D: priority queues: dense ranks (one wave per list), then one
    // thread per list pushes all pairs and pops the first POPP (all of them
    // after a round whose greedy popped the heaps often)
    {
        const int K = S.K;
        const int nq = S.size1 ? 1 : (K - 1) * (K - 1);
        for (int q = w; q < nq; q += NW) {
            int a, b;
            if (S.size1) { a = 0; b = S.self_pair ? 0 : 1; }
            else q_pair(q, K, a, b);
            int64_t key = 0;
            if (lane < MM) {
                int lg, rg;
                pair_gain(S, K, M, a, b, lane, lg, rg);
                key = swap_key(lg, rg);
                S.wkeys[w][lane] = key;
            }
            __builtin_amdgcn_wave_barrier();
            int rnk = 0;
            if (lane < MM) {
                // independent broadcast reads: unrolled so they issue back to
                // back instead of one LDS round trip per key
#pragma unroll 16
                for (int j = 0; j < MM; ++j) rnk += S.wkeys[w][j] < key;
            }
            // the push order is element order: stage entry e in slot e
            if (lane < MM) {
                HeapRef Hq{reinterpret_cast<uint16_t *>(S.heap[q]), 0};
                hset(Hq, lane, (uint32_t)((rnk << 8) | lane));
            }
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        if (tid == 0) st->clk[11] += __builtin_amdgcn_s_memrealtime() - tD;
        // one thread per list, all in wave 0: the heap steps are bound by
        // instruction issue, so 64 lists share every instruction
        const int npre = min(st->npre, MM);
        if (tid == 0) {
            S.npre = npre;
            S.heap_pops = 0;
            S.maxk = 0;
        }
        if (tid < nq) {
            HeapRef H{reinterpret_cast<uint16_t *>(S.heap[tid]), 0};
            for (int e = 0; e < MM; ++e) heap_push(H, hget(H, e));
            for (int k = 0; k < npre; ++k) S.popped[k][tid] = (uint8_t)(heap_pop(H) & 0xFFu);
            S.heapn[tid] = H.n;
            S.qcur[tid] = 0;
        }
        __syncthreads();
    }

    const uint64_t tE = __builtin_amdgcn_s_memrealtime();
    // ---- phase E: the sequential greedy (wave 0 in lockstep)
    if (w == 0) {
        const int nsw = M == 8 ? greedy_apply<8>(S, rev, drain) : greedy_apply<4>(S, rev, drain);
        if (lane == 0) {
            S.nsw = nsw;
            st->clk[12] += nsw;           // swaps
            st->clk[13] += S.heap_pops;   // pops taken from heaps by the greedy
            st->clk[14] += S.npre;        // pops precomputed per ordered pair
            // the next round precomputes what this one used, plus 8: a
            // round that had to pop heaps precomputes whole pop orders next
            st->npre = S.heap_pops > 0 ? MM : max(POPP, min(MM, (S.maxk + 15) / 8 * 8));
            const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
            st->clk[6] += tE - tD;
            st->clk[7] += t1 - tE;
        }
    }
    __syncthreads();
    if (S.status != 0) {
        if (tid == 0) {
            st->status = S.status;
            st->done = 1;
        }
        return;
    }

    // ---- phase F: commit the swaps to perm/inv (reorder(), :392-418)
    if (tid < S.nsw) {
        int v1 = S.swv1[tid], v2 = S.swv2[tid];
        if (v1 != v2) {
            int b1 = ld_l2(&A.inv[v1]), b2 = ld_l2(&A.inv[v2]);
            st_l2(&A.inv[v1], b2);
            st_l2(&A.inv[v2], b1);
            st_l2(&A.perm[b1], v2);
            st_l2(&A.perm[b2], v1);
        }
        if (st->trace) {
            int64_t n = st->trace_len + tid;
            if (n < st->trace_cap) {
                int32_t *t = st->trace + n * 9;
                t[0] = st->round; t[1] = S.swp[tid]; t[2] = S.swt[tid];
                t[3] = v1; t[4] = v2; t[5] = S.swlg[tid]; t[6] = S.swrg[tid];
                t[7] = S.swpc[tid]; t[8] = S.swtc[tid];
            }
        }
    }
    __syncthreads();
    if (tid == 0) st->trace_len += S.nsw;


    __syncthreads();
    // columns of the round's segments after the commit, for the recount
    const int K = S.K;
    const int total = table_columns(S.sel, K, M, A.inv, A.cptr, S.cs.cols, S.cs.pre);
    if (tid == 0) S.cs.total = total;
    chunk_table(S.cs, K, M);
    copy_words(&st->C, &S.cs, sizeof(ColSet));
    if (tid < KMAX) {
        st->newcnt[tid] = 0;
        st->newnzr[tid] = 0;
    }
    if (tid == 0) {
        st->pending = 1;
        st->work += total;
    }
    clock_add(st, 3, t0);
}

// Start of a run: parameters, initial list size and count
// (reorder/reorder.h:1808-1819).
__global__ __launch_bounds__(GT) void k_gr_init(NMArgs A, int R, int M, int N, int nseg, int sched,
                                                int maxiter, int round_cap, int32_t *trace,
                                                int64_t trace_cap) {
    __shared__ int sL, sC;
    GreedyState *st = A.st;
    const int tid = threadIdx.x;
    if (tid == 0) { sL = 0; sC = 0; }
    __syncthreads();
    int lsum = 0, csum = 0;
    for (int s = tid; s < nseg; s += GT) {
        const int c = A.cnt[s];
        lsum += c != 0;
        csum += c;
    }
    if (lsum) atomicAdd(&sL, lsum);
    if (csum) atomicAdd(&sC, csum);
    for (int t = tid; t < NACC; t += GT) st->acc[t] = 0;
    __syncthreads();
    if (tid == 0) {
        st->R = R; st->M = M; st->N = N; st->nseg = nseg; st->sched = sched;
        st->round_cap = round_cap; st->trace = trace; st->trace_cap = trace_cap;
        st->round = 1; st->L = sL; st->totcnt = sC;
        st->init_size = sL; st->init_cnt = sC;
        st->maxiter = maxiter > sL ? maxiter : sL;
        st->status = 0; st->done = 0; st->pending = 0; st->npre = POPP;
        st->K = 0; st->self_pair = 0; st->size1 = 0;
        st->trace_len = 0; st->work = 0;
        st->S.K = 0; st->S.total = 0; st->S.cb[0] = 0;
        st->C.K = 0; st->C.total = 0; st->C.cb[0] = 0;
        for (int i = 0; i < 16; ++i) st->clk[i] = 0;
    }
}

// ------------------------------------------------------------ host side

static int grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65535) g = 65535;
    return (int)g;
}

int ensure_tab(vnm_graph *g, size_t slots) {
    if (g->tab_slots >= slots && g->tab.p) return VNM_OK;
    size_t words = slots * (size_t)(g->nrows >> 2);
    g->tab.release();
    int rc = g->tab.alloc(words);
    if (rc) return rc;
    VNM_HIP_TRY(hipMemsetAsync(g->tab.p, 0, words * 4, g->ctx->stream));
    g->tab_slots = slots;
    return VNM_OK;
}

int graph_scratch(vnm_graph *g) {
    const int R = g->nrows;
    int rc;
    if (!g->cptr.p && (rc = g->cptr.alloc(R + 2))) return rc;
    if (!g->crow.p && (rc = g->crow.alloc(std::max<int64_t>(g->nnz, 1)))) return rc;
    if (!g->perm.p && (rc = g->perm.alloc(R))) return rc;
    if (!g->inv.p && (rc = g->inv.alloc(R))) return rc;
    if (!g->cnt.p && (rc = g->cnt.alloc(g->nseg() + 1))) return rc;
    if (!g->nzr.p && (rc = g->nzr.alloc(g->nseg() + 1))) return rc;
    size_t ws = std::max<size_t>(NPAIR_T * 2 * 64 + 64, (size_t)R / SCAN_TILE + 64 + R + 2);
    if (!g->ws.p && (rc = g->ws.alloc(ws))) return rc;
    return VNM_OK;
}

static int scan_inclusive(vnm_graph *g, int32_t *p, int64_t n, int32_t *tmp) {
    hipStream_t st = g->ctx->stream;
    int ntiles = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_TILE), 0, st, p, n, tmp);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(SCAN_TILE), 0, st, tmp, ntiles);
    hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(SCAN_TILE), 0, st, p, n, tmp);
    g->ctx->launches += 3;
    VNM_LAUNCH_CHECK("scan");
    return VNM_OK;
}

// CSC of the nonzero entries over current ids; perm = inv = identity.
int build_csc(vnm_graph *g) {
    int rc = graph_scratch(g);
    if (rc) return rc;
    hipStream_t st = g->ctx->stream;
    const int R = g->nrows;
    VNM_HIP_TRY(hipMemsetAsync(g->cptr.p, 0, (R + 1) * 4, st));
    hipLaunchKernelGGL(k_col_hist, dim3(grid_for(g->nnz, 256)), dim3(256), 0, st, g->row.p,
                       g->col.p, g->val.p, g->nnz, R, g->cptr.p);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_col_hist");
    rc = scan_inclusive(g, g->cptr.p, R + 1, g->ws.p);
    if (rc) return rc;
    int32_t *cursor = g->ws.p + (R / SCAN_TILE + 64);
    VNM_HIP_TRY(hipMemcpyAsync(cursor, g->cptr.p, (R + 1) * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_col_scatter, dim3(grid_for(g->nnz, 256)), dim3(256), 0, st, g->row.p,
                       g->col.p, g->val.p, g->nnz, R, cursor, g->crow.p);
    hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(R, 256)), dim3(256), 0, st, g->perm.p, (int64_t)R);
    hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(R, 256)), dim3(256), 0, st, g->inv.p, (int64_t)R);
    g->ctx->launches += 3;
    VNM_LAUNCH_CHECK("k_col_scatter");
    return VNM_OK;
}

// Keys (column << 32 | row) of the entries build_csc takes; the others get
// the largest key and sort last.
__global__ void k_csc_keys(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                           const float *__restrict__ val, int64_t nnz, int32_t R, uint64_t *__restrict__ keys) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = col[i], r = row[i];
        keys[i] = val[i] != 0.0f && c < R && r < R ? ((uint64_t)(uint32_t)c << 32) | (uint32_t)r : ~0ull;
    }
}

// First occurrence of each sorted key: its row, a flag, and a count for its
// column.
__global__ void k_csc_unique(const uint64_t *__restrict__ keys, int64_t n, int32_t *__restrict__ rows,
                             uint8_t *__restrict__ flags, int32_t *__restrict__ cnt1) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        const bool f = k != ~0ull && (i == 0 || keys[i - 1] != k);
        rows[i] = (int32_t)(uint32_t)k;
        flags[i] = f;
        if (f) atomicAdd(&cnt1[(int)(k >> 32) + 1], 1);
    }
}

// build_csc for a COO that may hold one position more than once (after the
// allpairs schedules' literal relabel, which is not a permutation): each
// position enters the CSC once, as a bitmap holds it once.  Rows within a
// column come out in ascending order.
int build_csc_unique(vnm_graph *g) {
    int rc = graph_scratch(g);
    if (rc) return rc;
    hipStream_t st = g->ctx->stream;
    const int R = g->nrows;
    const int64_t n = g->nnz;
    DBuf<uint64_t> k0, k1;
    DBuf<int32_t> rows, cnt;
    DBuf<uint8_t> flags;
    const size_t n1 = (size_t)std::max<int64_t>(n, 1);
    if ((rc = k0.alloc(n1)) || (rc = k1.alloc(n1)) || (rc = rows.alloc(n1)) || (rc = flags.alloc(n1)) ||
        (rc = cnt.alloc(1)))
        return rc;
    VNM_HIP_TRY(hipMemsetAsync(g->cptr.p, 0, (R + 1) * 4, st));
    if (n > 0) {
        hipLaunchKernelGGL(k_csc_keys, dim3(grid_for(n, 256)), dim3(256), 0, st, g->row.p, g->col.p, g->val.p, n, R,
                           k0.p);
        g->ctx->launches++;
        VNM_LAUNCH_CHECK("k_csc_keys");
        if ((rc = sort_keys_u64(k0.p, k1.p, n, 64, st))) return rc;
        hipLaunchKernelGGL(k_csc_unique, dim3(grid_for(n, 256)), dim3(256), 0, st, k1.p, n, rows.p, flags.p,
                           g->cptr.p);
        g->ctx->launches++;
        VNM_LAUNCH_CHECK("k_csc_unique");
        if ((rc = select_flagged_i32(rows.p, flags.p, g->crow.p, cnt.p, n, st))) return rc;
    }
    rc = scan_inclusive(g, g->cptr.p, R + 1, g->ws.p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(R, 256)), dim3(256), 0, st, g->perm.p, (int64_t)R);
    hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(R, 256)), dim3(256), 0, st, g->inv.p, (int64_t)R);
    g->ctx->launches += 2;
    VNM_LAUNCH_CHECK("k_iota_i32");
    // the scratch is freed on return: finish with it first
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

static int locate_grid(vnm_graph *g) { return std::max(1, std::min(g->nseg(), 1024)); }

int locate_all(vnm_graph *g, int N, int32_t *d_cnt, int32_t *d_nzr) {
    int grid = locate_grid(g);
    int rc = ensure_tab(g, std::max<size_t>(grid, KMAX + 1));
    if (rc) return rc;
    if (!g->queue.p && (rc = g->queue.alloc(4))) return rc;
    VNM_HIP_TRY(hipMemsetAsync(g->queue.p, 0, 4, g->ctx->stream));
    KernelTimer kt(g->ctx, "k_locate_all");
    hipLaunchKernelGGL(k_locate_all, dim3(grid), dim3(LOC_T), 0, g->ctx->stream, g->nrows, g->M, N,
                       g->nseg(), g->inv.p, g->cptr.p, g->crow.p, g->tab.p, d_cnt, d_nzr, g->queue.p);
    kt.stop();
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_locate_all");
    return VNM_OK;
}

int max_ids(vnm_graph *g, const int32_t *row, const int32_t *col, int32_t *max_row, int32_t *max_col) {
    hipStream_t st = g->ctx->stream;
    int32_t h[2] = {-1, -1};
    if (g->nnz > 0) {
        DBuf<int32_t> out;
        int rc = out.alloc(2);
        if (rc) return rc;
        VNM_HIP_TRY(hipMemcpyAsync(out.p, h, 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_max_ids, dim3(std::min<int64_t>(grid_for(g->nnz, 256), 1024)), dim3(256), 0, st,
                           row, col, g->nnz, out.p);
        g->ctx->launches++;
        VNM_LAUNCH_CHECK("k_max_ids");
        VNM_HIP_TRY(hipMemcpyAsync(h, out.p, 8, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
    }
    *max_row = h[0];
    *max_col = h[1];
    return VNM_OK;
}

int relabel(vnm_graph *g, const int32_t *d_perm) {
    hipStream_t st = g->ctx->stream;
    g->epoch++;
    KernelTimer kt(g->ctx, "k_relabel");
    hipLaunchKernelGGL(k_relabel, dim3(grid_for(g->nnz, 256)), dim3(256), 0, st, g->row.p,
                       g->col.p, g->nnz, d_perm, g->nrows);
    kt.stop();
    hipLaunchKernelGGL(k_compose, dim3(grid_for((int64_t)g->perm_total.n, 256)), dim3(256), 0, st,
                       g->perm_total.p, (int64_t)g->perm_total.n, d_perm, g->nrows);
    g->ctx->launches += 2;
    VNM_LAUNCH_CHECK("k_relabel");
    // the Reorder kernel renames columns with the row permutation
    // (reorder/reorder.h:23-35): with fewer columns than rows an entry can
    // move past the padded column count, which the per-segment passes of
    // VENOM and the packer must refuse
    const int M = g->M;
    if ((g->ncols + M - 1) / M * M < g->nrows && !g->col_overflow) {
        int32_t mr, mc;
        int rc = max_ids(g, g->row.p, g->col.p, &mr, &mc);
        if (rc) return rc;
        g->col_overflow = mc >= (g->ncols + M - 1) / M * M;
    }
    return VNM_OK;
}

static int check_nm_args(vnm_graph *g, int N) {
    if (g->M != 4 && g->M != 8)
        return set_error(VNM_E_ARG, "the GPU N:M engine supports M = 4 or 8 (got %d)", g->M);
    if (N < 1 || N >= g->M) return set_error(VNM_E_ARG, "N must be in [1, M) (got %d)", N);
    if (g->nrows % g->M != 0 || g->ncols > g->nrows)
        return set_error(VNM_E_ARG, "N:M reorder needs nrows %% M == 0 and ncols <= nrows");
    return VNM_OK;
}

constexpr int ROUND_BATCH = 16;  // greedy rounds per captured graph launch
constexpr int MAX_BATCH = 4096;  // graphs per batched run

// 1 in *flag when every graph of the batch has finished.
__global__ __launch_bounds__(64) void k_gr_alldone(const NMArgs *__restrict__ args, int G,
                                                   int32_t *__restrict__ flag) {
    int open = 0;
    for (int i = threadIdx.x; i < G; i += 64) open |= args[i].st->done == 0;
    const uint64_t b = __ballot(open);
    if (threadIdx.x == 0) *flag = b == 0;
}

// Capture ROUND_BATCH rounds (six kernels each, grid y = the G graphs) and
// the all-done test on the context stream.
static int capture_rounds(hipStream_t st, const NMArgs &a0, const NMArgs *d_args, int G, int gpx,
                          int32_t *d_flag, hipGraph_t *graph_out, hipGraphExec_t *exec_out) {
    VNM_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int b = 0; b < ROUND_BATCH; ++b) {
        hipLaunchKernelGGL(k_gr_select, dim3(1, G), dim3(GT), 0, st, a0, d_args);
        hipLaunchKernelGGL(k_gr_fill_a, dim3(gpx, G), dim3(PT), 0, st, a0, d_args);
        hipLaunchKernelGGL(k_gr_score, dim3(gpx, G), dim3(PT), 0, st, a0, d_args);
        hipLaunchKernelGGL(k_gr_greedy, dim3(1, G), dim3(GT), 0, st, a0, d_args);
        hipLaunchKernelGGL(k_gr_fill_b, dim3(gpx, G), dim3(PT), 0, st, a0, d_args);
        hipLaunchKernelGGL(k_gr_count, dim3(gpx, G), dim3(PT), 0, st, a0, d_args);
    }
    hipLaunchKernelGGL(k_gr_alldone, dim3(1), dim3(64), 0, st, d_args, G, d_flag);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_error(VNM_E_HIP, "capture of the N:M round kernels failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_error(VNM_E_HIP, "instantiating the N:M round graph failed: %s", hipGetErrorString(e));
    }
    *graph_out = graph;
    *exec_out = exec;
    return VNM_OK;
}

// Workgroups per graph of the record pass kernels: two per CU for up to four
// graphs (one graph's round can touch millions of entries), fewer per graph
// in larger batches (the grid-stride loops cover any count).
static int pass_grid(vnm_ctx *ctx, int G) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu < 1)
        ncu = 1;
    return std::max(16, std::min(2 * ncu, 8 * ncu / G));
}

// Bytes of the greedy state of a graph, then its one-graph argument block
// and all-done flag.
constexpr size_t GST_STATE = (sizeof(GreedyState) + 255) / 256 * 256;
constexpr size_t GST_BYTES = GST_STATE + (sizeof(NMArgs) + 63) / 64 * 64 + 64;

// Per-graph state of a run: CSC, counts, records, greedy state and the
// kernel arguments.
static int greedy_prepare(vnm_graph *g, int N, NMArgs *A) {
    int rc;
    if ((rc = build_csc(g))) return rc;
    if ((rc = locate_all(g, N, g->cnt.p, g->nzr.p))) return rc;
    hipStream_t st = g->ctx->stream;
    if (!g->gst.p && (rc = g->gst.alloc(GST_BYTES / 4))) return rc;
    if (!g->rt.p) {
        // records A and B, 16 bytes per row each, zero between runs
        if ((rc = g->rt.alloc((size_t)g->nrows * 8))) return rc;
        VNM_HIP_TRY(hipMemsetAsync(g->rt.p, 0, (size_t)g->nrows * 32, st));
    }
    A->st = (GreedyState *)g->gst.p;
    A->cptr = g->cptr.p; A->crow = g->crow.p; A->perm = g->perm.p; A->inv = g->inv.p;
    A->cnt = g->cnt.p; A->nzr = g->nzr.p;
    A->rta = g->rt.p; A->rtb = g->rt.p + (size_t)g->nrows * 4;
    return VNM_OK;
}

static void debug_print(int i, const GreedyState &h) {
    fprintf(stderr, "[vnm] greedy%s%d rounds=%d entries=%lld kernel_us select=%llu fill_a=%llu score=%llu "
                    "greedy=%llu (heaps %llu, sequential %llu) fill_b=%llu count=%llu\n",
            i < 0 ? "" : " graph ", i < 0 ? 0 : i,
            h.round - 1, (long long)h.work, (unsigned long long)(h.clk[0] / 100),
            (unsigned long long)(h.clk[1] / 100), (unsigned long long)(h.clk[2] / 100),
            (unsigned long long)(h.clk[3] / 100), (unsigned long long)(h.clk[6] / 100),
            (unsigned long long)(h.clk[7] / 100), (unsigned long long)(h.clk[4] / 100),
            (unsigned long long)(h.clk[5] / 100));
    fprintf(stderr, "[vnm]   select: commit %llu, segments %llu, columns %llu; heaps: ranks %llu\n",
            (unsigned long long)(h.clk[8] / 100), (unsigned long long)(h.clk[9] / 100),
            (unsigned long long)(h.clk[10] / 100), (unsigned long long)(h.clk[11] / 100));
    fprintf(stderr, "[vnm]   greedy: swaps %llu, heap pops %llu, precomputed pops per pair (sum) %llu\n",
            (unsigned long long)h.clk[12], (unsigned long long)h.clk[13], (unsigned long long)h.clk[14]);
}

// The greedy of G graphs (schedules 0-3 or >= 8 each) as one chain of round
// kernels on the context stream.  Graphs that finish early return at entry
// of the later rounds' kernels.  The trace is recorded only for G == 1.
static int greedy_run(vnm_graph *const *gs, const int *scheds, int G, int N, int maxiter, int round_cap,
                      int32_t *d_trace, int64_t trace_cap, int64_t *trace_len, vnm_nm_stats *stats) {
    vnm_ctx *ctx = gs[0]->ctx;
    hipStream_t st = ctx->stream;
    int rc;
    std::vector<NMArgs> hargs(G);
    for (int i = 0; i < G; ++i)
        if ((rc = greedy_prepare(gs[i], N, &hargs[i]))) return rc;
    // the argument array: inside the graph's state block for one graph (its
    // captured rounds are kept with the graph), a buffer of the call for many
    DBuf<int32_t> batch_buf;
    NMArgs *d_args;
    int32_t *d_flag;
    if (G == 1) {
        d_args = (NMArgs *)((char *)gs[0]->gst.p + GST_STATE);
        d_flag = (int32_t *)((char *)gs[0]->gst.p + GST_BYTES - 64);
    } else {
        const size_t abytes = ((size_t)G * sizeof(NMArgs) + 63) / 64 * 64;
        if ((rc = batch_buf.alloc((abytes + 64) / 4))) return rc;
        d_args = (NMArgs *)batch_buf.p;
        d_flag = (int32_t *)((char *)batch_buf.p + abytes);
    }
    VNM_HIP_TRY(hipMemcpyAsync(d_args, hargs.data(), (size_t)G * sizeof(NMArgs), hipMemcpyHostToDevice, st));
    const int gpx = pass_grid(ctx, G);
    KernelTimer kt(ctx, "nm_greedy");
    for (int i = 0; i < G; ++i) {
        vnm_graph *g = gs[i];
        hipLaunchKernelGGL(k_gr_init, dim3(1), dim3(GT), 0, st, hargs[i], g->nrows, g->M, N, g->nseg(),
                           scheds[i], maxiter, round_cap > 0 ? round_cap : 0x7FFFFFFF, d_trace,
                           d_trace ? trace_cap : (int64_t)0);
    }
    ctx->launches += G;
    VNM_LAUNCH_CHECK("k_gr_init");
    // rounds run in captured batches of ROUND_BATCH; the kernels of rounds
    // past a graph's end return at entry
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (G == 1) {
        // the captured rounds hold the arguments by value: capture again when
        // they changed (a reload that shrinks the rows moves records B)
        vnm_graph *g0 = gs[0];
        const uint8_t *key = (const uint8_t *)&hargs[0];
        if (g0->nm_exec && (g0->nm_key.size() != sizeof(NMArgs) ||
                            memcmp(g0->nm_key.data(), key, sizeof(NMArgs)) != 0)) {
            (void)hipGraphExecDestroy(g0->nm_exec);
            (void)hipGraphDestroy(g0->nm_graph);
            g0->nm_exec = nullptr;
            g0->nm_graph = nullptr;
        }
        if (!g0->nm_exec) {
            if ((rc = capture_rounds(st, hargs[0], d_args, 1, gpx, d_flag, &g0->nm_graph, &g0->nm_exec)))
                return rc;
            g0->nm_key.assign(key, key + sizeof(NMArgs));
        }
        exec = g0->nm_exec;
    } else if ((rc = capture_rounds(st, hargs[0], d_args, G, gpx, d_flag, &graph, &exec))) {
        return rc;
    }
    struct Drop {
        hipGraph_t g;
        hipGraphExec_t e;
        ~Drop() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } drop{graph, G == 1 ? nullptr : exec};
    int32_t done = 0;
    while (!done) {
        VNM_HIP_TRY(hipGraphLaunch(exec, st));
        ctx->launches += 6 * ROUND_BATCH + 1;
        VNM_HIP_TRY(hipMemcpyAsync(&done, d_flag, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
    }
    for (int i = 0; i < G; ++i)
        hipLaunchKernelGGL(k_gr_cleanup, dim3(gpx), dim3(PT), 0, st, hargs[i], d_args + i);
    ctx->launches += G;
    VNM_LAUNCH_CHECK("k_gr_cleanup");
    kt.stop();
    std::vector<GreedyState> h(G);
    for (int i = 0; i < G; ++i)
        VNM_HIP_TRY(hipMemcpyAsync(&h[i], hargs[i].st, sizeof(GreedyState), hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    const bool debug = getenv("VNM_DEBUG") != nullptr;
    int ub = -1;
    for (int i = 0; i < G; ++i) {
        if (debug) debug_print(G == 1 ? -1 : i, h[i]);
        if (h[i].status == VNM_E_UB_PATH) {
            if (ub < 0) ub = i;
            continue;
        }
        if ((rc = relabel(gs[i], gs[i]->perm.p))) return rc;
        stats[i].init_size = h[i].init_size;
        stats[i].final_size = h[i].L;
        stats[i].init_cnt = h[i].init_cnt;
        stats[i].final_cnt = h[i].totcnt;
        stats[i].rounds = h[i].round - 1;
    }
    if (ub >= 0)
        return set_error(VNM_E_UB_PATH, "greedy of graph %d reached top() of an empty swap queue "
                                        "(undefined in the reference)", ub);
    if (trace_len) *trace_len = h[0].trace_len;
    return VNM_OK;
}

int nm_reorder_ex(vnm_graph *g, int N, int sched, int maxiter, int round_cap, int32_t *d_trace,
                  int64_t trace_cap, int64_t *trace_len, vnm_nm_stats *stats) {
    int rc = check_nm_args(g, N);
    if (rc) return rc;
    if (sched >= 4 && sched <= 7) {
        if (d_trace || round_cap > 0) return set_error(VNM_E_ARG, "allpairs schedules record no trace");
        return nm_reorder_allpairs(g, N, sched, maxiter, 0, stats);
    }
    if (sched < 0) return set_error(VNM_E_ARG, "schedule %d is negative", sched);
    return greedy_run(&g, &sched, 1, N, maxiter, round_cap, d_trace, trace_cap, trace_len, stats);
}

// The greedy of many graphs in one launch chain.  Each graph keeps its own
// schedule, stats and result; they are those of nm_reorder on the graph alone.
int nm_reorder_batch(vnm_graph *const *gs, int count, int N, const int *scheds, int maxiter,
                     vnm_nm_stats *stats) {
    if (count < 1 || count > MAX_BATCH) return set_error(VNM_E_ARG, "batch of %d graphs (1..%d)", count, MAX_BATCH);
    for (int i = 0; i < count; ++i) {
        if (!gs[i]) return set_error(VNM_E_ARG, "graph %d is null", i);
        if (gs[i]->ctx != gs[0]->ctx) return set_error(VNM_E_ARG, "graphs of a batch share one context");
        int rc = check_nm_args(gs[i], N);
        if (rc) return rc;
        if (scheds[i] < 0 || (scheds[i] >= 4 && scheds[i] <= 7))
            return set_error(VNM_E_ARG, "batched schedules are 0-3 or >= 8 (graph %d: %d)", i, scheds[i]);
        for (int j = 0; j < i; ++j)
            if (gs[j] == gs[i]) return set_error(VNM_E_ARG, "graph %d appears twice in the batch", i);
    }
    return greedy_run(gs, scheds, count, N, maxiter, 0, nullptr, 0, nullptr, stats);
}

int nm_reorder(vnm_graph *g, int N, int sched, int maxiter, vnm_nm_stats *stats) {
    return nm_reorder_ex(g, N, sched, maxiter, 0, nullptr, 0, nullptr, stats);
}

}  // namespace vnm

using namespace vnm;

extern "C" {

int vnm_locate(vnm_graph *g, int N, int32_t *cnt, int32_t *zero) {
    if (!g) return set_error(VNM_E_ARG, "null graph");
    int rc = check_nm_args(g, N);
    if (rc) return rc;
    if ((rc = build_csc(g))) return rc;
    if ((rc = locate_all(g, N, g->cnt.p, g->nzr.p))) return rc;
    const int ns = g->nseg();
    std::vector<int32_t> nz(ns);
    hipStream_t st = g->ctx->stream;
    if (cnt) VNM_HIP_TRY(hipMemcpyAsync(cnt, g->cnt.p, ns * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipMemcpyAsync(nz.data(), g->nzr.p, ns * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    if (zero)
        for (int s = 0; s < ns; ++s) zero[s] = g->nrows - nz[s];
    return VNM_OK;
}

int vnm_swap_score(vnm_graph *g, int N, int seg1, const int32_t *seg2, int count, int32_t *scores) {
    if (!g || count < 0 || (count > 0 && (!seg2 || !scores))) return set_error(VNM_E_ARG, "bad arguments");
    int rc = check_nm_args(g, N);
    if (rc) return rc;
    const int ns = g->nseg(), M = g->M;
    if (seg1 < 0 || seg1 >= ns) return set_error(VNM_E_ARG, "seg1 out of range");
    for (int k = 0; k < count; ++k)
        if (seg2[k] < 0 || seg2[k] >= ns) return set_error(VNM_E_ARG, "seg2[%d] out of range", k);
    if (count == 0) return VNM_OK;
    if ((rc = build_csc(g))) return rc;
    hipStream_t st = g->ctx->stream;
    const int chunk = 64;
    if ((rc = ensure_tab(g, std::max<size_t>(chunk + 1, locate_grid(g))))) return rc;
    DBuf<int32_t> dseg, dsc;
    if ((rc = dseg.alloc(chunk + 1)) || (rc = dsc.alloc((size_t)chunk * 2 * M * M))) return rc;
    for (int base = 0; base < count; base += chunk) {
        int n = std::min(chunk, count - base);
        std::vector<int32_t> segs(n + 1);
        segs[0] = seg1;
        for (int k = 0; k < n; ++k) segs[k + 1] = seg2[base + k];
        VNM_HIP_TRY(hipMemcpyAsync(dseg.p, segs.data(), (n + 1) * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemsetAsync(dsc.p, 0, (size_t)n * 2 * M * M * 4, st));
        hipLaunchKernelGGL(k_fill_slots, dim3(n + 1), dim3(256), 0, st, g->nrows, M, dseg.p,
                           g->inv.p, g->cptr.p, g->crow.p, g->tab.p, 0);
        int gx = std::max(1, std::min((g->nrows + SC_ROWS - 1) / SC_ROWS, 1024));
        KernelTimer kt(g->ctx, "k_score_rows");
        hipLaunchKernelGGL(k_score_rows, dim3(gx, n), dim3(SC_ROWS), 0, st, g->nrows, M, N, n,
                           g->tab.p, dsc.p);
        kt.stop();
        hipLaunchKernelGGL(k_fill_slots, dim3(n + 1), dim3(256), 0, st, g->nrows, M, dseg.p,
                           g->inv.p, g->cptr.p, g->crow.p, g->tab.p, 1);
        g->ctx->launches += 3;
        VNM_LAUNCH_CHECK("k_score_rows");
        VNM_HIP_TRY(hipMemcpyAsync(scores + (size_t)base * 2 * M * M, dsc.p,
                                   (size_t)n * 2 * M * M * 4, hipMemcpyDeviceToHost, st));
    }
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int vnm_pq_pop_order(vnm_ctx *ctx, const int32_t *scores, int count, int M, int32_t *order) {
    if (!ctx || count < 0 || M < 1 || M > 8) return set_error(VNM_E_ARG, "bad arguments");
    if (count == 0) return VNM_OK;
    const int MM = M * M;
    DBuf<int32_t> ds, dord;
    int rc;
    if ((rc = ds.alloc((size_t)count * 2 * MM)) || (rc = dord.alloc((size_t)count * MM))) return rc;
    hipStream_t st = ctx->stream;
    VNM_HIP_TRY(hipMemcpyAsync(ds.p, scores, (size_t)count * 2 * MM * 4, hipMemcpyHostToDevice, st));
    const int bt = 64;
    hipLaunchKernelGGL(k_pq_pop_order, dim3((count + bt - 1) / bt), dim3(bt), (size_t)bt * HEAP_WORDS * 4, st, ds.p,
                       count, M, dord.p);
    ctx->launches++;
    VNM_LAUNCH_CHECK("k_pq_pop_order");
    VNM_HIP_TRY(hipMemcpyAsync(order, dord.p, (size_t)count * MM * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int vnm_nm_reorder(vnm_graph *g, int N, int sched, int maxiter, vnm_nm_stats *st) {
    if (!g || !st) return set_error(VNM_E_ARG, "bad arguments");
    return nm_reorder(g, N, sched, maxiter, st);
}

int vnm_nm_reorder_batch(vnm_graph *const *graphs, int count, int N, const int32_t *sched, int maxiter,
                         vnm_nm_stats *st) {
    if (!graphs || !sched || !st) return set_error(VNM_E_ARG, "bad arguments");
    return nm_reorder_batch(graphs, count, N, sched, maxiter, st);
}

int vnm_nm_trace(vnm_graph *g, int N, int sched, int maxiter, int round_cap, int32_t *rec,
                 int64_t cap, int64_t *count, vnm_nm_stats *st) {
    if (!g || !st || cap < 0 || (cap > 0 && !rec)) return set_error(VNM_E_ARG, "bad arguments");
    if (cap > ((int64_t)1 << 31) / 9) return set_error(VNM_E_ARG, "trace capacity too large");
    DBuf<int32_t> d;
    int rc;
    if ((rc = d.alloc((size_t)std::max<int64_t>(cap, 1) * 9))) return rc;
    int64_t n = 0;
    if ((rc = nm_reorder_ex(g, N, sched, maxiter, round_cap, d.p, cap, &n, st))) return rc;
    int64_t keep = std::min(n, cap);
    if (keep > 0) {
        VNM_HIP_TRY(hipMemcpyAsync(rec, d.p, (size_t)keep * 9 * 4, hipMemcpyDeviceToHost, g->ctx->stream));
        VNM_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    }
    if (count) *count = n;
    return VNM_OK;
}

}  // extern "C"
