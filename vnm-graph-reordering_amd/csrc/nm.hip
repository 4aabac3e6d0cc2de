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
//   k_nm_greedy   : the whole greedy schedule in one persistent workgroup
//   k_relabel     : the Reorder kernel (reorder/reorder.h:23-35), applied once
#include "graph.h"
#include "hip_common.h"
#include "tables.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
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
    int32_t *__restrict__ cnt, int32_t *__restrict__ nzr) {
    __shared__ ColSpan cols[16];
    __shared__ int pre[17];
    __shared__ int sel[1];
    __shared__ int c_inv, c_nz;
    uint32_t *slot = tab + (size_t)blockIdx.x * (R >> 2);
    for (int s = blockIdx.x; s < nseg; s += gridDim.x) {
        if (threadIdx.x == 0) { sel[0] = s; c_inv = 0; c_nz = 0; }
        __syncthreads();
        int total = table_columns(sel, 1, M, inv, cptr, cols, pre);
        table_scatter(slot, nullptr, M, R, cols, pre, M, total, crow);
        __syncthreads();
        int my_inv = 0, my_nz = 0;
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            int t = entry_col(pre, M, e);
            int r = crow[cols[t].start + (e - pre[t])];
            uint32_t bits = slot_byte(slot, 0, R, r);
            if ((bits >> (M - 1 - t)) == 1u) {  // leading bit: the row's owner
                my_nz += 1;
                my_inv += popc32(bits) > N;
            }
        }
        if (my_inv) atomicAdd(&c_inv, my_inv);
        if (my_nz) atomicAdd(&c_nz, my_nz);
        __syncthreads();
        table_clear(slot, nullptr, M, R, cols, pre, M, total, crow);
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
    if (clear) table_clear(slot, nullptr, M, R, cols, pre, M, total, crow);
    else table_scatter(slot, nullptr, M, R, cols, pre, M, total, crow);
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

__global__ void k_pq_pop_order(const int32_t *__restrict__ scores, int count, int M,
                               int32_t *__restrict__ order) {
    extern __shared__ unsigned char smem[];
    const int MM = M * M;
    // slot-major layout: heap slot s of thread t at [s * blockDim.x + t]
    uint8_t *ranks = smem;
    uint8_t *heaps = smem + (size_t)blockDim.x * MM;
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int32_t *sc = scores + (size_t)q * 2 * MM;
    uint8_t *rank = ranks + threadIdx.x;
    for (int i = 0; i < MM; ++i) {
        int64_t ki = swap_key(sc[2 * i], sc[2 * i + 1]);
        int r = 0;
        for (int j = 0; j < MM; ++j) r += swap_key(sc[2 * j], sc[2 * j + 1]) < ki;
        rank[i * blockDim.x] = (uint8_t)r;
    }
    HeapRef H{heaps + threadIdx.x, rank, 0, (int)blockDim.x};
    for (int i = 0; i < MM; ++i) heap_push(H, i);
    for (int k = 0; k < MM; ++k) order[(size_t)q * MM + k] = heap_pop(H);
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

__global__ void k_compose(int32_t *__restrict__ total, const int32_t *__restrict__ p, int32_t n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) total[i] = p[total[i]];
}

// ------------------------------------------------------------ greedy

constexpr int GT = 1024;         // threads of the persistent workgroup
constexpr int NW = GT / 64;      // its waves
constexpr int KMAX = 9;          // primary + up to M targets (M <= 8)
constexpr int MMAX = 8;
constexpr int NPAIR_T = KMAX * (KMAX - 1) / 2 + 1;  // unordered pairs + one self pair
constexpr int NQ = (KMAX - 1) * (KMAX - 1);         // ordered pairs the greedy can visit
constexpr int POPP = 16;                            // pops precomputed per ordered pair

struct NMArgs {
    int R, M, N, nseg, sched, maxiter;
    const int32_t *cptr, *crow;
    int32_t *perm, *inv, *cnt, *nzr;
    uint32_t *tab;       // KMAX slots of R bytes, zero on entry and exit
    int32_t *out;        // stats[5], status, work, nswaps
    int32_t *trace;      // optional: 9 ints per swap
    int64_t trace_cap;
    int round_cap;       // stop after this many rounds (training traces; INT_MAX = none)
};

struct GreedyShared {
    int sel[KMAX];
    int selcnt[KMAX];
    int K, self_pair, size1, cont;
    int L, totcnt, round, maxiter, init_size, init_cnt, status;
    ColSpan cols[KMAX * MMAX];
    int pre[KMAX * MMAX + 1];
    int cb[KMAX + 1];
    int E[KMAX][MMAX];
    // per unordered pair (j < j2) and one self pair: the "both rows" sums
    //   Bl[b1*M+b2] = sum over rows in both of y_b2 * (x_b1 ? A : B)
    //   Br[b1*M+b2] = sum over rows in both of x_b1 * (y_b2 ? D : C)
    int both[NPAIR_T][2][MMAX * MMAX];
    int newcnt[KMAX], newnzr[KMAX];
    // slot-major (element / slot / pop index first, list second), so the
    // lanes of a wave building different lists touch consecutive bytes
    uint8_t rank[MMAX * MMAX][NQ];
    uint8_t heap[MMAX * MMAX][NQ];
    uint8_t popped[POPP][NQ];
    int heapn[NQ];
    uint8_t vis[KMAX * MMAX];
    int swv1[64], swv2[64], swlg[64], swrg[64], swp[64], swt[64], swpc[64], swtc[64];
    int nsw;
    uint64_t wtop[NW][KMAX + 1];
    uint64_t wbest[NW];
    uint64_t topk[KMAX + 1];
    int64_t wkeys[NW][64];
    uint64_t phase_clk[8];
    int64_t trace_len;
    int64_t work;
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

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t lo = __shfl((uint32_t)v, src, 64), hi = __shfl((uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
    const int lane = threadIdx.x & 63;
    for (int d = 32; d >= 1; d >>= 1) {
        uint64_t o = shfl_u64(v, lane ^ d);
        v = o > v ? o : v;
    }
    return v;
}

// Top-k (k <= KMAX + 1) keys of the block into S.topk (0 = none); tk holds
// each thread's descending local list.
__device__ void block_topk(uint64_t (&tk)[KMAX + 1], int k, GreedyShared &S) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int it = 0; it < k; ++it) {
        uint64_t m = wave_max_u64(tk[0]);
        if (m != 0 && tk[0] == m) {
#pragma unroll
            for (int i = 0; i < KMAX; ++i) tk[i] = tk[i + 1];
            tk[KMAX] = 0;
        }
        if (lane == 0) S.wtop[w][it] = m;
    }
    __syncthreads();
    if (w == 0) {
        uint64_t lst[KMAX + 1];
#pragma unroll
        for (int i = 0; i <= KMAX; ++i) lst[i] = (lane < NW && i < k) ? S.wtop[lane][i] : 0ull;
        for (int it = 0; it < k; ++it) {
            uint64_t m = wave_max_u64(lst[0]);
            if (m != 0 && lst[0] == m) {
#pragma unroll
                for (int i = 0; i < KMAX; ++i) lst[i] = lst[i + 1];
                lst[KMAX] = 0;
            }
            if (lane == 0) S.topk[it] = m;
        }
    }
    __syncthreads();
}

__device__ uint64_t block_max_u64(uint64_t v, GreedyShared &S) {
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

__device__ void q_pair(int q, int K, int &a, int &b) {
    const int n = K - 1;
    if (q < n) { a = 0; b = q + 1; return; }
    q -= n;
    a = 1 + q / (n - 1);
    int r = q % (n - 1);
    b = 1 + r + (r + 1 >= a ? 1 : 0);
}

// Decode entry e of slot j: its column bit b and row r.
__device__ __forceinline__ void slot_entry(const GreedyShared &S, const int32_t *crow, int j, int M,
                                           int e, int &b, int &r) {
    b = entry_col(S.pre + j * M, M, e);
    const int t = j * M + b;
    r = crow[S.cols[t].start + (e - S.pre[t])];
}

// Chunk table of the round: slot j owns 64-entry chunks [cb[j], cb[j+1]).
__device__ int chunk_table(GreedyShared &S, int K, int M) {
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int j = 0; j < K; ++j) {
            S.cb[j] = acc;
            acc += (S.pre[j * M + M] - S.pre[j * M] + 63) / 64;
        }
        S.cb[K] = acc;
    }
    __syncthreads();
    return S.cb[K];
}

__device__ __forceinline__ int chunk_slot(const GreedyShared &S, int K, int c) {
    int j = 0;
    while (j + 1 < K && S.cb[j + 1] <= c) ++j;
    return j;
}

// Gains of every pair of the round's segments (the GetSegSwapScore sums,
// reorder/reorder.h:75-131, in closed form).  Waves take 64-entry chunks of
// one slot each; the owner entries of a chunk are folded in with ballots, so
// each (b1, b2) counter takes one LDS atomic per chunk instead of one per row,
// and the partner-slot bytes of a chunk are loaded together.
__device__ void greedy_score_phase(const NMArgs &A, GreedyShared &S) {
    const int M = A.M, N = A.N, R = A.R, K = S.K, MM = M * M;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nc = K * M;
    for (int t = threadIdx.x; t < KMAX * MMAX; t += GT) (&S.E[0][0])[t] = 0;
    for (int t = threadIdx.x; t < NPAIR_T * 2 * MMAX * MMAX; t += GT) (&S.both[0][0][0])[t] = 0;
    const int total = table_columns(S.sel, K, M, A.inv, A.cptr, S.cols, S.pre);
    table_scatter(A.tab, nullptr, M, R, S.cols, S.pre, nc, total, A.crow);
    const int nch = chunk_table(S, K, M);  // (barrier inside)
    const int b1 = lane / M, b2 = lane % M;
    for (int c = w; c < nch; c += NW) {
        const int j = chunk_slot(S, K, c);
        const int e1 = S.pre[j * M + M];
        const int e = S.pre[j * M] + (c - S.cb[j]) * 64 + lane;
        uint32_t l = 0;
        int r = 0;
        bool act = false;
        if (e < e1) {
            int b;
            slot_entry(S, A.crow, j, M, e, b, r);
            l = slot_byte(A.tab, j, R, r);
            act = (l >> (M - 1 - b)) == 1u;  // the row's leading-bit entry
            if (!act) l = 0;
        }
        uint32_t rrs[KMAX];
#pragma unroll
        for (int j2 = 0; j2 < KMAX; ++j2)
            rrs[j2] = (act && j2 > j && j2 < K) ? slot_byte(A.tab, j2, R, r) : 0u;
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
            if (v) atomicAdd(&S.E[j][lane], v);
        }
#pragma unroll
        for (int j2 = 0; j2 < KMAX; ++j2) {
            if (j2 <= j || j2 >= K) continue;
            const uint32_t rr = rrs[j2];
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
                if (vl) atomicAdd(&S.both[pi][0][lane], vl);
                if (vr) atomicAdd(&S.both[pi][1][lane], vr);
            }
        }
        if (S.self_pair && lane < MM) {  // seg1 == seg2: r = l for every row
            const int vl = __popcll(Lb2 & ((Lb1 & Am) | (~Lb1 & Bm)));
            const int vr = __popcll(Lb1 & ((Lb2 & Am) | (~Lb2 & Bm)));
            if (vl) atomicAdd(&S.both[NPAIR_T - 1][0][lane], vl);
            if (vr) atomicAdd(&S.both[NPAIR_T - 1][1][lane], vr);
        }
    }
    __syncthreads();
    table_clear(A.tab, nullptr, M, R, S.cols, S.pre, nc, total, A.crow);
    if (threadIdx.x == 0) S.work += total;
    __syncthreads();
}

// cnt / nzr of the round's segments after the swaps (locate_invalid_segment).
__device__ void greedy_recount_phase(const NMArgs &A, GreedyShared &S) {
    const int M = A.M, N = A.N, R = A.R, K = S.K;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nc = K * M;
    if (threadIdx.x < KMAX) { S.newcnt[threadIdx.x] = 0; S.newnzr[threadIdx.x] = 0; }
    const int total = table_columns(S.sel, K, M, A.inv, A.cptr, S.cols, S.pre);
    table_scatter(A.tab, nullptr, M, R, S.cols, S.pre, nc, total, A.crow);
    const int nch = chunk_table(S, K, M);
    for (int c = w; c < nch; c += NW) {
        const int j = chunk_slot(S, K, c);
        const int e1 = S.pre[j * M + M];
        const int e = S.pre[j * M] + (c - S.cb[j]) * 64 + lane;
        bool act = false;
        uint32_t l = 0;
        if (e < e1) {
            int b, r;
            slot_entry(S, A.crow, j, M, e, b, r);
            l = slot_byte(A.tab, j, R, r);
            act = (l >> (M - 1 - b)) == 1u;
        }
        const uint64_t o = __ballot(act), v = __ballot(act && popc32(l) > N);
        if (lane == 0) {
            if (o) atomicAdd(&S.newnzr[j], (int)__popcll(o));
            if (v) atomicAdd(&S.newcnt[j], (int)__popcll(v));
        }
    }
    __syncthreads();
    table_clear(A.tab, nullptr, M, R, S.cols, S.pre, nc, total, A.crow);
    if (threadIdx.x == 0) S.work += total;
    __syncthreads();
}

// First element of ordered pair q (slots a -> b) in pop order whose two
// vertices are unvisited; -1 when the queue runs empty.
__device__ int first_free(GreedyShared &S, int q, int a, int b, int M) {
    const int npre = min(POPP, M * M);
    for (int k = 0; k < npre; ++k) {
        int e = S.popped[k][q];
        if (!S.vis[a * M + e / M] && !S.vis[b * M + e % M]) return e;
    }
    HeapRef H{&S.heap[0][q], &S.rank[0][q], S.heapn[q], NQ};
    while (H.n > 0) {
        int e = heap_pop(H);
        if (!S.vis[a * M + e / M] && !S.vis[b * M + e % M]) { S.heapn[q] = H.n; return e; }
    }
    S.heapn[q] = 0;
    return -1;
}

__device__ __forceinline__ bool list_before(int sa, int ca, int sb, int cb, bool rev) {
    // true when (sa, ca) is nearer the top of the invalid list than (sb, cb)
    if (!rev) return ca != cb ? ca > cb : sa < sb;
    return ca != cb ? ca < cb : sa > sb;
}

// The sequential part of one round (reorder/reorder.h:1837-1932), one thread.
__device__ void greedy_apply(GreedyShared &S, int M, bool rev, bool drain) {
    const int K = S.K;
    S.nsw = 0;
    for (int i = 0; i < K * M; ++i) S.vis[i] = 0;
    auto add_swap = [&](int a, int b, int e, int lg, int rg, int pc, int tc) {
        int n = S.nsw++;
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
        return;
    }
    int cur = 0, pc = S.selcnt[0];
    int tgt[KMAX], tcnt[KMAX], nt = K - 1;
    for (int i = 0; i < nt; ++i) { tgt[i] = i + 1; tcnt[i] = S.selcnt[i + 1]; }
    while (true) {
        int upd[KMAX], ucnt[KMAX], nu = 0;
        for (int i = 0; i < nt; ++i) {
            bool all = true;
            for (int v = 0; v < M; ++v) all = all && S.vis[cur * M + v];
            if (pc <= 0 || all) break;  // reorder/reorder.h:1878
            const int t = tgt[i];
            const int q = q_index(cur, t, K);
            int e = first_free(S, q, cur, t, M);
            if (e < 0) {
                // top() of an empty queue: the reference would read the last
                // popped pair, which has a visited endpoint (:1908)
                S.status = VNM_E_UB_PATH;
                return;
            }
            int lg, rg;
            pair_gain(S, K, M, cur, t, e, lg, rg);
            S.vis[cur * M + e / M] = 1;
            S.vis[t * M + e % M] = 1;
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
}

// Phase clock (wall_clock64 ticks, 100 MHz on gfx950) accumulated per phase
// for the diagnostic counters returned in out[8..15].
#define PHASE_MARK(i)                                                     \
    do {                                                                  \
        if (tid == 0) {                                                   \
            uint64_t _t = __builtin_amdgcn_s_memrealtime();               \
            S.phase_clk[i] += _t - t_last;                                \
            t_last = _t;                                                  \
        }                                                                 \
    } while (0)

__global__ __launch_bounds__(GT) void k_nm_greedy(NMArgs A) {
    __shared__ GreedyShared S;
    uint64_t t_last = __builtin_amdgcn_s_memrealtime();
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int M = A.M, R = A.R, nseg = A.nseg, MM = M * M;
    const bool plain = A.sched >= 8;
    const bool rev = (A.sched == 2 || A.sched == 3);
    const bool drain = (A.sched == 1 || A.sched == 3);
    const bool tfwd = drain ? rev : !rev;  // traversal order uses the forward key
    const int min_size = plain ? 2 : 1;

    // initial list size and count (reorder/reorder.h:1808-1810)
    if (tid == 0) {
        S.L = 0; S.totcnt = 0; S.work = 0;
        for (int i = 0; i < 8; ++i) S.phase_clk[i] = 0;
    }
    __syncthreads();
    {
        int lsum = 0, csum = 0;
        for (int s = tid; s < nseg; s += GT) {
            int c = A.cnt[s];
            lsum += c != 0;
            csum += c;
        }
        if (lsum) atomicAdd(&S.L, lsum);
        if (csum) atomicAdd(&S.totcnt, csum);
    }
    __syncthreads();
    if (tid == 0) {
        S.init_size = S.L;
        S.init_cnt = S.totcnt;
        S.maxiter = A.maxiter > S.L ? A.maxiter : S.L;  // :1819
        S.round = 1;
        S.status = 0;
        S.trace_len = 0;
    }
    __syncthreads();

    while (true) {
        if (tid == 0)
            S.cont = (S.round <= S.maxiter && S.round <= A.round_cap && S.L >= min_size && S.status == 0);
        __syncthreads();
        if (!S.cont) break;

        // ---- phase A: the round's segments
        if (!plain && S.L == 1) {
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
            const int k = min(M, S.L - 1) + 1;  // targets + the primary itself
            uint64_t tk[KMAX + 1];
#pragma unroll
            for (int i = 0; i <= KMAX; ++i) tk[i] = 0;
            uint64_t best_l = 0;
            for (int s = tid; s < nseg; s += GT) {
                int c = ld_l2(&A.cnt[s]);
                if (c == 0) continue;
                uint64_t lk = rev ? key_rev(s, c) : key_fwd(s, c);
                best_l = lk > best_l ? lk : best_l;
                uint64_t x = tfwd ? key_fwd(s, c) : key_rev(s, c);
#pragma unroll
                for (int i = 0; i <= KMAX; ++i) {
                    uint64_t hi = tk[i] > x ? tk[i] : x;
                    x = tk[i] > x ? x : tk[i];
                    tk[i] = hi;
                }
            }
            uint64_t bl = block_max_u64(best_l, S);
            block_topk(tk, k, S);
            if (tid == 0) {
                int primary = seg_of_key(bl, !rev);
                S.size1 = 0;
                S.self_pair = 0;
                S.sel[0] = primary;
                S.selcnt[0] = ld_l2(&A.cnt[primary]);
                int K = 1;
                for (int i = 0; i < k && K < k; ++i) {
                    if (S.topk[i] == 0) break;
                    int s = seg_of_key(S.topk[i], tfwd);
                    if (s == primary) continue;
                    S.sel[K] = s;
                    S.selcnt[K] = ld_l2(&A.cnt[s]);
                    ++K;
                }
                S.K = K;
            }
            __syncthreads();
        }

        PHASE_MARK(0);
        // ---- phase C: gains of every pair of the round's segments
        greedy_score_phase(A, S);
        PHASE_MARK(1);

        // ---- phase D: priority queues: dense ranks (one wave per list), then
        // one thread per list pushes all pairs and pops the first POPP
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
                if (lane < MM)
                    for (int j = 0; j < MM; ++j) rnk += S.wkeys[w][j] < key;
                if (lane < MM) S.rank[lane][q] = (uint8_t)rnk;
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
            if (tid < nq) {
                HeapRef H{&S.heap[0][tid], &S.rank[0][tid], 0, NQ};
                for (int e = 0; e < MM; ++e) heap_push(H, e);
                const int npre = min(POPP, MM);
                for (int k = 0; k < npre; ++k) S.popped[k][tid] = (uint8_t)heap_pop(H);
                S.heapn[tid] = H.n;
            }
            __syncthreads();
        }

        PHASE_MARK(2);
        // ---- phase E: the sequential greedy (one thread)
        if (tid == 0) greedy_apply(S, M, rev, drain);
        __syncthreads();
        PHASE_MARK(3);
        if (S.status != 0) break;

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
            if (A.trace) {
                int64_t n = S.trace_len + tid;
                if (n < A.trace_cap) {
                    int32_t *t = A.trace + n * 9;
                    t[0] = S.round; t[1] = S.swp[tid]; t[2] = S.swt[tid];
                    t[3] = v1; t[4] = v2; t[5] = S.swlg[tid]; t[6] = S.swrg[tid];
                    t[7] = S.swpc[tid]; t[8] = S.swtc[tid];
                }
            }
        }
        __syncthreads();
        if (tid == 0) S.trace_len += S.nsw;
        __syncthreads();

        PHASE_MARK(4);
        // ---- phase G: recount the touched segments (locate_invalid_segment)
        greedy_recount_phase(A, S);
        PHASE_MARK(5);
        if (tid == 0) {
            for (int j = 0; j < S.K; ++j) {
                int s = S.sel[j];
                int oc = ld_l2(&A.cnt[s]);
                int nc = S.newcnt[j];
                st_l2(&A.cnt[s], nc);
                st_l2(&A.nzr[s], S.newnzr[j]);
                S.L += (nc != 0) - (oc != 0);
                S.totcnt += nc - oc;
            }
            S.round += 1;
        }
        __syncthreads();
    }
    if (tid == 0) {
        A.out[0] = S.init_size;
        A.out[1] = S.L;
        A.out[2] = S.init_cnt;
        A.out[3] = S.totcnt;
        A.out[4] = S.round - 1;
        A.out[5] = S.status;
        A.out[6] = (int32_t)min(S.work, (int64_t)0x7FFFFFFF);
        A.out[7] = (int32_t)S.trace_len;
        for (int i = 0; i < 6; ++i) A.out[8 + i] = (int32_t)min(S.phase_clk[i] / 100, (uint64_t)0x7FFFFFFF);
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

static int locate_grid(vnm_graph *g) { return std::max(1, std::min(g->nseg(), 1024)); }

int locate_all(vnm_graph *g, int N, int32_t *d_cnt, int32_t *d_nzr) {
    int grid = locate_grid(g);
    int rc = ensure_tab(g, std::max<size_t>(grid, KMAX + 1));
    if (rc) return rc;
    KernelTimer kt(g->ctx, "k_locate_all");
    hipLaunchKernelGGL(k_locate_all, dim3(grid), dim3(LOC_T), 0, g->ctx->stream, g->nrows, g->M, N,
                       g->nseg(), g->inv.p, g->cptr.p, g->crow.p, g->tab.p, d_cnt, d_nzr);
    kt.stop();
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_locate_all");
    return VNM_OK;
}

int relabel(vnm_graph *g, const int32_t *d_perm) {
    hipStream_t st = g->ctx->stream;
    g->epoch++;
    hipLaunchKernelGGL(k_relabel, dim3(grid_for(g->nnz, 256)), dim3(256), 0, st, g->row.p,
                       g->col.p, g->nnz, d_perm, g->nrows);
    hipLaunchKernelGGL(k_compose, dim3(grid_for(g->nrows, 256)), dim3(256), 0, st,
                       g->perm_total.p, d_perm, g->nrows);
    g->ctx->launches += 2;
    VNM_LAUNCH_CHECK("k_relabel");
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

int nm_reorder_ex(vnm_graph *g, int N, int sched, int maxiter, int round_cap, int32_t *d_trace,
                  int64_t trace_cap, int64_t *trace_len, vnm_nm_stats *stats) {
    int rc = check_nm_args(g, N);
    if (rc) return rc;
    if (sched >= 4 && sched <= 7) {
        if (d_trace || round_cap > 0) return set_error(VNM_E_ARG, "allpairs schedules record no trace");
        return nm_reorder_allpairs(g, N, sched, maxiter, 0, stats);
    }
    if (sched < 0) return set_error(VNM_E_ARG, "schedule %d is negative", sched);
    if ((rc = build_csc(g))) return rc;
    if ((rc = locate_all(g, N, g->cnt.p, g->nzr.p))) return rc;
    hipStream_t st = g->ctx->stream;
    int32_t *out = g->ws.p + NPAIR_T * 2 * 64;
    NMArgs A;
    A.R = g->nrows; A.M = g->M; A.N = N; A.nseg = g->nseg(); A.sched = sched; A.maxiter = maxiter;
    A.cptr = g->cptr.p; A.crow = g->crow.p; A.perm = g->perm.p; A.inv = g->inv.p;
    A.cnt = g->cnt.p; A.nzr = g->nzr.p; A.tab = g->tab.p; A.out = out;
    A.trace = d_trace; A.trace_cap = d_trace ? trace_cap : 0;
    A.round_cap = round_cap > 0 ? round_cap : 0x7FFFFFFF;
    KernelTimer kt(g->ctx, "k_nm_greedy");
    hipLaunchKernelGGL(k_nm_greedy, dim3(1), dim3(GT), 0, st, A);
    kt.stop();
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_nm_greedy");
    int32_t h[16];
    VNM_HIP_TRY(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    if (getenv("VNM_DEBUG"))
        fprintf(stderr, "[vnm] greedy rounds=%d entries=%d phase_us A=%d C=%d D=%d E=%d F=%d G=%d\n",
                h[4], h[6], h[8], h[9], h[10], h[11], h[12], h[13]);
    if (h[5] == VNM_E_UB_PATH)
        return set_error(VNM_E_UB_PATH, "greedy reached top() of an empty swap queue "
                                        "(undefined in the reference)");
    if ((rc = relabel(g, g->perm.p))) return rc;
    if (trace_len) *trace_len = h[7];
    stats->init_size = h[0];
    stats->final_size = h[1];
    stats->init_cnt = h[2];
    stats->final_cnt = h[3];
    stats->rounds = h[4];
    return VNM_OK;
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
    size_t shm = (size_t)bt * MM * 2;
    hipLaunchKernelGGL(k_pq_pop_order, dim3((count + bt - 1) / bt), dim3(bt), shm, st, ds.p, count,
                       M, dord.p);
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
