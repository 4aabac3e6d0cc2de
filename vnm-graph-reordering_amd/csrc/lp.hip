// Learned pair proposer for the N:M reorder (gfx950, wave64).
//
// An addition of this project: the reference has no learned component.  Its
// exact arithmetic -- the per-row swap gain of GetSegSwapScore
// (reorder/reorder.h:75-131) -- stays the judge: a graph network ranks
// candidate column pairs, the top pairs are scored exactly by the
// entry-based pair scorer below, and only swaps with a positive exact gain
// are applied, segment-disjoint within an iteration so that every applied
// gain is exact.
//
// Network (nodes = matrix ids, node u's list = rows r with a nonzero (r, u)):
//   features  16 per node from the class counts of u's rows in u's segment
//             (class of a row in a segment: popcount < N, = N, = N+1, > N+1)
//   layers    H_{l+1} = relu([H_l, Agg H_l] W_l + b_l), Agg = neighbour mean,
//             widths 16 -> 64 -> 64 -> 64
//   head      Y = H3 Wh + bh: Y0 donor score, Y1 acceptor score, Y2..17 = P,
//             Y18..33 = Q; pair score s(u, v) = Y0[u] + Y1[v] + P[u].Q[v]
//   training  squared error of s against the exact gain lgain + rgain
//             (target transform applied by the caller), Adam.
// BF16 storage of activations and weights with FP32 accumulation on the MFMA
// path; an FP32 VALU path computes the same network for gradient checks.
// Backward is FP32 throughout.
//
// Kernels:
//   k_lp_adj_keys/unpack : deterministic (sorted) adjacency and its transpose
//   k_lp_seg_stats       : class counts per node and per segment (whole graph
//                          or touched segments) from an LDS row hash or LDS
//                          row tiles, segments taken from a work counter
//   k_lp_features        : 16 input features per node
//   k_lp_agg_*           : neighbour mean (and its transpose for backward),
//                          one wave per node, long lists split into chunks
//                          summed in a fixed order (bit-reproducible)
//   k_lp_dense_mfma      : [H, Agg] W + b with v_mfma_f32_16x16x16_bf16
//   k_lp_dense_f32       : the same on VALU in FP32
//   k_lp_pair_loss, k_lp_bwd_in, k_lp_wgrad(+reduce), k_lp_adam
//   k_lp_pick/match      : donor and acceptor per segment, candidate pairs
//   k_lp_pair_gain       : exact M x M gain matrix of a segment pair from the
//                          entries of its two segments (ballot bit planes)
//   k_lp_select/swap     : segment-disjoint greedy over positive exact gains
//                          (parallel rounds, the sequential walk's choice)
#include "graph.h"
#include "hip_common.h"
#include "sortlib.h"
#include "tables.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

struct vnm_lp_model {
    vnm_ctx *ctx = nullptr;
    vnm::DBuf<float> w, m, v, g;   // FP32 master weights, Adam moments, gradient
    vnm::DBuf<uint16_t> wt;        // BF16 copies, each layer transposed [64][K]
    int64_t step = 0;
};

namespace vnm {

constexpr int LP_F = 16, LP_H = 64, LP_HD = 64, LP_PQ = 16;
constexpr int LP_NPARAMS = 22784;
constexpr int LP_OFF_W[3] = {0, 2112, 10368};
constexpr int LP_OFF_B[3] = {2048, 10304, 18560};
constexpr int LP_OFF_WH = 18624, LP_OFF_BH = 22720;
constexpr int LP_WT_OFF[4] = {0, 2048, 10240, 18432};  // BF16 [64][K] copies
constexpr int LP_NWT = 22528;
constexpr int AGG_CH = 1024;   // neighbours per aggregation chunk
constexpr int WG_GRID = 256;   // workgroups of the weight-gradient reduction

// ------------------------------------------------------------ storage types

template <typename T> struct Act;
template <> struct Act<float> {
    static __device__ __forceinline__ float4 load4(const float *p) { return *(const float4 *)p; }
    static __device__ __forceinline__ float get(const float *p) { return *p; }
    static __device__ __forceinline__ void put(float *p, float v) { *p = v; }
};
template <> struct Act<uint16_t> {
    static __device__ __forceinline__ float4 load4(const uint16_t *p) {
        uint2 w = *(const uint2 *)p;
        return make_float4(bf16_to_f32((uint16_t)(w.x & 0xFFFFu)), bf16_to_f32((uint16_t)(w.x >> 16)),
                           bf16_to_f32((uint16_t)(w.y & 0xFFFFu)), bf16_to_f32((uint16_t)(w.y >> 16)));
    }
    static __device__ __forceinline__ float get(const uint16_t *p) { return bf16_to_f32(*p); }
    static __device__ __forceinline__ void put(uint16_t *p, float v) { *p = f32_to_bf16(v); }
};

__device__ __forceinline__ uint32_t ord_f32(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint32_t mix32(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)((x ^ (x >> 31)) >> 32);
}

static int grid_of(int64_t n, int block, int cap = 65535) {
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// ------------------------------------------------------------ adjacency

// key = (list owner << 32) | member; entries that are not part of the graph
// get owner R so they sort after every real key.
__global__ void k_lp_adj_keys(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                              const float *__restrict__ val, int64_t nnz, int R,
                              uint64_t *__restrict__ kc, uint64_t *__restrict__ kt,
                              int32_t *__restrict__ nvalid) {
    int local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = row[i], c = col[i];
        const bool ok = val[i] != 0.0f && r >= 0 && r < R && c >= 0 && c < R;
        kc[i] = ok ? ((uint64_t)(uint32_t)c << 32) | (uint32_t)r : (uint64_t)(uint32_t)R << 32;
        kt[i] = ok ? ((uint64_t)(uint32_t)r << 32) | (uint32_t)c : (uint64_t)(uint32_t)R << 32;
        local += ok;
    }
    if (local) atomicAdd(nvalid, local);
}

// Sorted keys (node << 32 | neighbour), n of them valid: idx[i] = neighbour,
// ptr[u] = the first i whose node is >= u (ptr[R] = n), written at the
// boundaries where the node changes (no per-entry counting).
__global__ void k_lp_adj_unpack(const uint64_t *__restrict__ keys, int64_t n, int R,
                                int32_t *__restrict__ idx, int32_t *__restrict__ ptr) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        idx[i] = (int32_t)(uint32_t)k;
        const int u = (int)(k >> 32);
        const int prev = i == 0 ? -1 : (int)(keys[i - 1] >> 32);
        for (int v = prev + 1; v <= u; ++v) ptr[v] = (int32_t)i;
        if (i == n - 1)
            for (int v = u + 1; v <= R; ++v) ptr[v] = (int32_t)n;
    }
}

__global__ void k_lp_iota2(int n, int32_t *__restrict__ a, int32_t *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = i; b[i] = i; }
}

__global__ void k_lp_inv_deg(int R, const int32_t *__restrict__ ptr, float *__restrict__ inv_deg) {
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < R) inv_deg[u] = 1.0f / (float)max(ptr[u + 1] - ptr[u], 1);
}

// ------------------------------------------------------------ segment statistics

constexpr int SS_T = 256;
constexpr int TB = SS_T / 8;  // row tiles whose ends one pass of binary searches finds
constexpr int PB = 256 / 16;   // the same for the pair scorer (2M <= 16 columns)
constexpr int PAIR_TILE_ROWS = TILE_ROWS / 2;  // rows of a pair tile (16 bits per row)

// For each selected segment (all when segs == nullptr): class counts of the
// rows of each of its M nodes (ccls[node][4]) and of the segment itself
// (segcls[seg][4], every row counted once), and its entry count.
__global__ __launch_bounds__(SS_T) void k_lp_seg_stats(
    int R, int M, int N, int nseg, const int32_t *__restrict__ segs, int nsel,
    const int32_t *__restrict__ inv, const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
    int32_t *__restrict__ ccls, int32_t *__restrict__ segcls, int32_t *__restrict__ segnnz,
    int32_t *__restrict__ queue, const int *__restrict__ d_nsel, const int *__restrict__ skip) {
    // a device-side count and skip flag, when given, replace nsel
    if (skip && *skip) return;
    if (d_nsel) nsel = *d_nsel;
    __shared__ ColSpan cols[16];
    __shared__ int pre[17];
    __shared__ int sel[1], cur;
    __shared__ int colc[8][4];
    __shared__ int segc[4];
    __shared__ int tlo[8], thi[8], tpre[9];
    __shared__ int tend[8][TB];
    // the row bitmaps of one segment: an LDS hash (at most HASH_MAX entries)
    // or, for larger segments, LDS byte tiles of TILE_ROWS rows
    __shared__ union {
        RowHash h;
        uint32_t tile[TILE_ROWS / 4];
    } L;
    const int n = segs ? nsel : nseg;
    const int tid = threadIdx.x;
    hash_reset(L.h);
    // segments from a work counter (*queue zero at launch): segment sizes are
    // heavy-tailed and cluster by index, so a fixed assignment leaves one
    // workgroup with many of the largest
    for (;;) {
        if (tid == 0) cur = atomicAdd(queue, 1);
        __syncthreads();
        const int i = cur;
        if (i >= n) break;
        if (tid == 0) sel[0] = segs ? segs[i] : i;
        if (tid < 32) colc[tid >> 2][tid & 3] = 0;
        if (tid < 4) segc[tid] = 0;
        __syncthreads();
        const int total = table_columns(sel, 1, M, inv, ptr, cols, pre);
        auto count = [&](int t, uint32_t bits) {
            const int pc = popc32(bits);
            const int k = pc < N ? 0 : pc == N ? 1 : pc == N + 1 ? 2 : 3;
            atomicAdd(&colc[t][k], 1);
            if ((bits >> (M - 1 - t)) == 1u) atomicAdd(&segc[k], 1);
        };
        if (total <= HASH_MAX) {
            hash_scatter(L.h, M, cols, pre, M, total, idx);
            __syncthreads();
            // TAB_U entries per thread resolved before their loads are issued
            for (int e0 = tid; e0 < total; e0 += SS_T * TAB_U) {
                int t[TAB_U], r[TAB_U];
                table_entries(cols, pre, M, total, idx, e0, SS_T, t, r);
#pragma unroll
                for (int u = 0; u < TAB_U; ++u)
                    if (t[u] >= 0) count(t[u], hash_get(L.h, (uint32_t)r[u]) & 0xFFu);
            }
            __syncthreads();
            hash_reset(L.h);
        } else {
            // rows [r0, r0 + TILE_ROWS) at a time in an LDS byte tile: the
            // column lists are sorted by row, so each column's entries of a
            // tile are one range, and the tile ends of TB tiles are found by
            // M * TB binary searches at once
            for (int w = tid; w < TILE_ROWS / 4; w += SS_T) L.tile[w] = 0u;
            if (tid < M) tlo[tid] = cols[tid].start;
            int cc[8][4], sc[4];  // this thread's class counts (columns, segment)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                sc[k] = 0;
#pragma unroll
                for (int t = 0; t < 8; ++t) cc[t][k] = 0;
            }
            const int ntile = (R + TILE_ROWS - 1) / TILE_ROWS;
            for (int g0 = 0; g0 < ntile; g0 += TB) {
                if (tid < M * TB) {  // end of tile g0 + j in column t
                    const int t = tid / TB, j = tid % TB;
                    const int64_t r1 = min((int64_t)R, (int64_t)(g0 + j + 1) * TILE_ROWS);
                    int lo = cols[t].start, hi = cols[t].start + cols[t].len;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (idx[mid] < r1) lo = mid + 1;
                        else hi = mid;
                    }
                    tend[t][j] = lo;
                }
                __syncthreads();
                for (int j = 0; j < TB && g0 + j < ntile; ++j) {
                    const int r0 = (g0 + j) * TILE_ROWS;
                    if (tid == 0) {
                        tpre[0] = 0;
                        for (int t = 0; t < M; ++t) {
                            thi[t] = tend[t][j];
                            tpre[t + 1] = tpre[t] + thi[t] - tlo[t];
                        }
                    }
                    __syncthreads();
                    const int tt = tpre[M];
                    if (tt > 0) {  // uniform
                        for (int e0 = tid; e0 < tt; e0 += SS_T * TAB_U) {
                            int t[TAB_U], r[TAB_U];
#pragma unroll
                            for (int u = 0; u < TAB_U; ++u) {
                                const int e = e0 + u * SS_T;
                                t[u] = e < tt ? entry_col(tpre, M, e) : -1;
                            }
#pragma unroll
                            for (int u = 0; u < TAB_U; ++u)
                                r[u] = t[u] >= 0 ? idx[tlo[t[u]] + (e0 + u * SS_T - tpre[t[u]])] - r0 : 0;
#pragma unroll
                            for (int u = 0; u < TAB_U; ++u)
                                if (t[u] >= 0) atomicOr(&L.tile[r[u] >> 2], (1u << (M - 1 - t[u])) << (8 * (r[u] & 3)));
                        }
                        __syncthreads();
                        // every row of the tile once: its bitmap's class,
                        // counted for the segment and for each of its columns
                        for (int w = tid; w < TILE_ROWS / 4; w += SS_T) {
                            const uint32_t word = L.tile[w];
                            if (!word) continue;
                            L.tile[w] = 0u;
#pragma unroll
                            for (int y = 0; y < 4; ++y) {
                                const uint32_t bits = (word >> (8 * y)) & 0xFFu;
                                if (!bits) continue;
                                const int pc = popc32(bits);
                                const int kk = pc < N ? 0 : pc == N ? 1 : pc == N + 1 ? 2 : 3;
#pragma unroll
                                for (int k = 0; k < 4; ++k) {
                                    const int is = kk == k;
                                    sc[k] += is;
#pragma unroll
                                    for (int t = 0; t < 8; ++t)
                                        if (t < M) cc[t][k] += is & (int)((bits >> (M - 1 - t)) & 1u);
                                }
                            }
                        }
                    }
                    if (tid < M) tlo[tid] = thi[tid];
                    __syncthreads();
                }
            }
            // the threads' counts into the segment's counters
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int v = sc[k];
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                if ((tid & 63) == 0 && v) atomicAdd(&segc[k], v);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    int c = cc[t][k];
                    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
                    if ((tid & 63) == 0 && c) atomicAdd(&colc[t][k], c);
                }
            }
            __syncthreads();
            hash_reset(L.h);
        }
        const int s = sel[0];
        if (tid < M * 4) {
            const int node = ld_l2(&inv[s * M + (tid >> 2)]);
            ccls[node * 4 + (tid & 3)] = colc[tid >> 2][tid & 3];
        }
        if (tid < 4) segcls[s * 4 + tid] = segc[tid];
        if (tid == 0) segnnz[s] = total;
        __syncthreads();
    }
}

// tot[0] = sum over segments of rows with more than N ones, tot[1] = number
// of segments with at least one such row.
__global__ __launch_bounds__(1024) void k_lp_totals(int nseg, const int32_t *__restrict__ segcls,
                                                     int32_t *__restrict__ tot, const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ int sc[16], ss[16];
    int c = 0, s = 0;
    for (int i = threadIdx.x; i < nseg; i += 1024) {
        int v = segcls[i * 4 + 2] + segcls[i * 4 + 3];
        c += v;
        s += v != 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o);
        s += __shfl_xor(s, o);
    }
    if ((threadIdx.x & 63) == 0) { sc[threadIdx.x >> 6] = c; ss[threadIdx.x >> 6] = s; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int w = 0; w < 16; ++w) { a += sc[w]; b += ss[w]; }
        tot[0] = a;
        tot[1] = b;
    }
}

template <typename T>
__global__ void k_lp_features(int R, int M, const int32_t *__restrict__ perm,
                              const int32_t *__restrict__ ptr, const int32_t *__restrict__ ccls,
                              const int32_t *__restrict__ segcls, const int32_t *__restrict__ segnnz,
                              const int32_t *__restrict__ tot, T *__restrict__ X, const int *__restrict__ skip) {
    if (skip && *skip) return;  // not wanted this iteration (a device-side flag)

    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= R) return;
    const int p = perm[u] / M;
    const int deg = ptr[u + 1] - ptr[u];
    const int4 c = *(const int4 *)&ccls[u * 4];
    const int4 S = *(const int4 *)&segcls[p * 4];
    const float dd = (float)max(deg, 1);
    float x[LP_F];
    x[0] = 1.0f;
    x[1] = log1pf((float)deg);
    x[2] = c.x / dd;
    x[3] = c.y / dd;
    x[4] = c.z / dd;
    x[5] = c.w / dd;
    x[6] = log1pf((float)c.z);
    x[7] = log1pf((float)c.y);
    x[8] = log1pf((float)c.w);
    x[9] = log1pf((float)(S.z + S.w));
    x[10] = log1pf((float)S.y);
    x[11] = log1pf((float)S.z);
    x[12] = log1pf((float)segnnz[p]);
    x[13] = (S.z + S.w) > 0 ? 1.0f : 0.0f;
    x[14] = log1pf((float)(S.y - c.y));
    x[15] = log1pf((float)tot[0]) * (1.0f / 16.0f);
    T *o = X + (size_t)u * LP_F;
#pragma unroll
    for (int k = 0; k < LP_F; ++k) Act<T>::put(o + k, x[k]);
}

// ------------------------------------------------------------ aggregation

// Forward (BWD = false): out[u] = inv_deg[u] * sum_{j in list(u)} H[j].
// Transpose (BWD = true): out[u] += sum_{j in tlist(u)} inv_deg[j] * H[j].
// One wave per node: F/4 lanes cover a row (4 values each), 256/F rows in
// flight; the partial sums of the row slots are combined by a fixed xor tree.
template <typename TI, int F, bool BWD>
__device__ __forceinline__ float4 agg_range(int b, int e, const int32_t *__restrict__ idx,
                                            const float *__restrict__ inv_deg, const TI *__restrict__ H) {
    constexpr int LPR = F / 4, NS = 64 / LPR;
    const int lane = threadIdx.x & 63;
    const int gq = lane % LPR, slot = lane / LPR;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    // four rows per slot in flight; each slot still sums its rows in list
    // order, so the result does not depend on the unrolling
    constexpr int UN = 4;
    int j = b + slot;
    for (; j + (UN - 1) * NS < e; j += UN * NS) {
        int v[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) v[u] = idx[j + u * NS];
        float4 h[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) h[u] = Act<TI>::load4(H + (size_t)v[u] * F + gq * 4);
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            if (BWD) {
                const float f = inv_deg[v[u]];
                h[u].x *= f; h[u].y *= f; h[u].z *= f; h[u].w *= f;
            }
            acc.x += h[u].x; acc.y += h[u].y; acc.z += h[u].z; acc.w += h[u].w;
        }
    }
    for (; j < e; j += NS) {
        const int v = idx[j];
        float4 h = Act<TI>::load4(H + (size_t)v * F + gq * 4);
        if (BWD) {
            const float f = inv_deg[v];
            h.x *= f; h.y *= f; h.z *= f; h.w *= f;
        }
        acc.x += h.x; acc.y += h.y; acc.z += h.z; acc.w += h.w;
    }
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1) {
        acc.x += __shfl_xor(acc.x, off);
        acc.y += __shfl_xor(acc.y, off);
        acc.z += __shfl_xor(acc.z, off);
        acc.w += __shfl_xor(acc.w, off);
    }
    return acc;
}

template <typename TI, typename TO, int F, bool BWD>
__global__ __launch_bounds__(256) void k_lp_agg_nodes(int R, const int32_t *__restrict__ ptr,
                                                      const int32_t *__restrict__ idx,
                                                      const float *__restrict__ inv_deg,
                                                      const TI *__restrict__ H, TO *__restrict__ out, const int *__restrict__ skip) {
    if (skip && *skip) return;  // not wanted this iteration (a device-side flag)

    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= R) return;
    const int b = ptr[u], e = ptr[u + 1];
    if (e - b > AGG_CH) return;  // split into chunks
    const float4 acc = agg_range<TI, F, BWD>(b, e, idx, inv_deg, H);
    constexpr int LPR = F / 4;
    const int lane = threadIdx.x & 63;
    if (lane < LPR) {
        TO *o = out + (size_t)u * F + lane * 4;
        if (BWD) {
            o[0] += acc.x; o[1] += acc.y; o[2] += acc.z; o[3] += acc.w;
        } else {
            const float f = inv_deg[u];
            Act<TO>::put(o + 0, acc.x * f);
            Act<TO>::put(o + 1, acc.y * f);
            Act<TO>::put(o + 2, acc.z * f);
            Act<TO>::put(o + 3, acc.w * f);
        }
    }
}

template <typename TI, int F, bool BWD>
__global__ __launch_bounds__(256) void k_lp_agg_chunks(int nch, const int32_t *__restrict__ ch_beg,
                                                       const int32_t *__restrict__ ch_end,
                                                       const int32_t *__restrict__ idx,
                                                       const float *__restrict__ inv_deg,
                                                       const TI *__restrict__ H, float *__restrict__ part, const int *__restrict__ skip) {
    if (skip && *skip) return;  // not wanted this iteration (a device-side flag)

    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nch) return;
    const float4 acc = agg_range<TI, F, BWD>(ch_beg[c], ch_end[c], idx, inv_deg, H);
    const int lane = threadIdx.x & 63;
    if (lane < F / 4) *(float4 *)&part[(size_t)c * F + lane * 4] = acc;
}

template <typename TO, int F, bool BWD>
__global__ void k_lp_agg_combine(int nheavy, const int32_t *__restrict__ heavy,
                                 const int32_t *__restrict__ ch0, const float *__restrict__ part,
                                 const float *__restrict__ inv_deg, TO *__restrict__ out, const int *__restrict__ skip) {
    if (skip && *skip) return;  // not wanted this iteration (a device-side flag)

    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nheavy * F) return;
    const int h = t / F, f = t % F;
    float s = 0.f;
    for (int c = ch0[h]; c < ch0[h + 1]; ++c) s += part[(size_t)c * F + f];
    const int u = heavy[h];
    TO *o = out + (size_t)u * F + f;
    if (BWD) *o += s;
    else Act<TO>::put(o, s * inv_deg[u]);
}

// ------------------------------------------------------------ dense layers

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// out[R x 64] = act([A | B] W + bias), A: R x KA, B: R x KB (BF16), W given
// transposed as Wt[64][KA+KB] (BF16).  One wave per 16 rows: for each 16-wide
// k step the wave's A fragment (lane l: row l%16, k 4*(l/16)..+3) is
// multiplied by four 16x16 B fragments (lane l: column l%16 of the n tile),
// accumulating a 16 x 64 FP32 tile (lane l: rows 4*(l/16)..+3, column l%16).
template <int KA, int KB, bool RELU, typename TO>
__global__ __launch_bounds__(256) void k_lp_dense_mfma(int R, const uint16_t *__restrict__ A,
                                                       const uint16_t *__restrict__ B,
                                                       const uint16_t *__restrict__ Wt,
                                                       const float *__restrict__ bias,
                                                       TO *__restrict__ out, const int *__restrict__ skip) {
    if (skip && *skip) return;  // not wanted this iteration (a device-side flag)

    constexpr int K = KA + KB, KP = K + 4;  // padded LDS row (bank spread)
    __shared__ uint16_t w[64 * KP];
    __shared__ float bs[64];
    for (int i = threadIdx.x; i < 64 * (K / 4); i += 256) {
        const int n = i / (K / 4), k4 = i % (K / 4);
        *(uint2 *)&w[n * KP + k4 * 4] = *(const uint2 *)&Wt[n * K + k4 * 4];
    }
    if (threadIdx.x < 64) bs[threadIdx.x] = bias[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + wv) * 16;
    if (r0 >= R) return;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    const int row = r0 + m;
    f32x4_t acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
        const int k = ks * 16 + kq;
        bf16x4_t a = bf16x4_t{0, 0, 0, 0};
        if (row < R) {
            const uint16_t *src = (k < KA) ? A + (size_t)row * KA + k : B + (size_t)row * KB + (k - KA);
            a = *(const bf16x4_t *)src;
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x4_t b = *(const bf16x4_t *)&w[(nt * 16 + m) * KP + k];
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int col = nt * 16 + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int rr = r0 + kq + i;
            if (rr < R) {
                float v = acc[nt][i] + bs[col];
                if (RELU) v = fmaxf(v, 0.f);
                Act<TO>::put(out + (size_t)rr * 64 + col, v);
            }
        }
    }
}

// The same layer in FP32 on VALU, W row-major [K][64].
template <int KA, int KB, bool RELU>
__global__ __launch_bounds__(256) void k_lp_dense_f32(int R, const float *__restrict__ A,
                                                      const float *__restrict__ B,
                                                      const float *__restrict__ W,
                                                      const float *__restrict__ bias,
                                                      float *__restrict__ out, const int *__restrict__ skip) {
    if (skip && *skip) return;  // not wanted this iteration (a device-side flag)

    constexpr int K = KA + KB;
    __shared__ float w[K * 64];
    for (int i = threadIdx.x; i < K * 64; i += 256) w[i] = W[i];
    __syncthreads();
    const int col = threadIdx.x & 63;
    const float b0 = bias[col];
    for (int row = blockIdx.x * 4 + (threadIdx.x >> 6); row < R; row += gridDim.x * 4) {
        float s = b0;
        for (int k = 0; k < KA; ++k) s += A[(size_t)row * KA + k] * w[k * 64 + col];
        for (int k = 0; k < KB; ++k) s += B[(size_t)row * KB + k] * w[(KA + k) * 64 + col];
        if (RELU) s = fmaxf(s, 0.f);
        out[(size_t)row * 64 + col] = s;
    }
}

// ------------------------------------------------------------ loss and backward

__global__ __launch_bounds__(256) void k_lp_pair_loss(int count, const int32_t *__restrict__ pairs,
                                                      const float *__restrict__ target,
                                                      const float *__restrict__ Y, float *__restrict__ dY,
                                                      float *__restrict__ part) {
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < count) {
        const int u = pairs[2 * i], v = pairs[2 * i + 1];
        const float *yu = Y + (size_t)u * LP_HD, *yv = Y + (size_t)v * LP_HD;
        float s = yu[0] + yv[1];
        for (int j = 0; j < LP_PQ; ++j) s += yu[2 + j] * yv[2 + LP_PQ + j];
        const float d = s - target[i];
        l = d * d;
        const float ds = 2.0f * d / (float)count;
        atomicAdd(&dY[(size_t)u * LP_HD + 0], ds);
        atomicAdd(&dY[(size_t)v * LP_HD + 1], ds);
        for (int j = 0; j < LP_PQ; ++j) {
            atomicAdd(&dY[(size_t)u * LP_HD + 2 + j], ds * yv[2 + LP_PQ + j]);
            atomicAdd(&dY[(size_t)v * LP_HD + 2 + LP_PQ + j], ds * yu[2 + j]);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dA = dPre W[:KA]^T, dB = dPre W[KA:]^T with dPre = dOut * [Hout > 0] (MASK)
// or dOut.  Four rows per step; W staged transposed so lanes read
// consecutive k.
template <typename T, int KA, int KB, bool MASK>
__global__ __launch_bounds__(256) void k_lp_bwd_in(int R, const float *__restrict__ dOut,
                                                   const T *__restrict__ Hout,
                                                   const float *__restrict__ W,
                                                   float *__restrict__ dA, float *__restrict__ dB) {
    constexpr int K = KA + KB;
    __shared__ float wt[64 * K];
    __shared__ float d[4][64];
    for (int i = threadIdx.x; i < K * 64; i += 256) {
        const int k = i / 64, n = i % 64;
        wt[n * K + k] = W[i];
    }
    for (int r0 = blockIdx.x * 4; r0 < R; r0 += gridDim.x * 4) {
        __syncthreads();
        {
            const int q = threadIdx.x >> 6, n = threadIdx.x & 63, rr = r0 + q;
            float gv = 0.f;
            if (rr < R) {
                gv = dOut[(size_t)rr * 64 + n];
                if (MASK && !(Act<T>::get(Hout + (size_t)rr * 64 + n) > 0.f)) gv = 0.f;
            }
            d[q][n] = gv;
        }
        __syncthreads();
        for (int o = threadIdx.x; o < 4 * K; o += 256) {
            const int q = o / K, k = o % K, rr = r0 + q;
            if (rr >= R) continue;
            float s = 0.f;
#pragma unroll 8
            for (int n = 0; n < 64; ++n) s += d[q][n] * wt[n * K + k];
            if (k < KA) dA[(size_t)rr * KA + k] = s;
            else dB[(size_t)rr * KB + (k - KA)] = s;
        }
    }
}

// Partial weight and bias gradients of one layer over a strided share of
// the rows: part[blk][k*64 + n] = sum_r [A|B][r][k] dPre[r][n], then the 64
// bias sums.  k_lp_wgrad_reduce adds the partials in block order.
template <typename T, int KA, int KB, bool MASK>
__global__ __launch_bounds__(256) void k_lp_wgrad(int R, const T *__restrict__ A, const T *__restrict__ B,
                                                  const float *__restrict__ dOut,
                                                  const T *__restrict__ Hout, float *__restrict__ part) {
    // thread (kb, n) owns outputs k = kb + 4j of column n, j < K/4, and, for
    // kb == 0, the bias of column n
    constexpr int K = KA + KB, RW = 16, NOUT = K * 64 + 64, NJ = K / 4;
    __shared__ float cin[RW][K];
    __shared__ float d[RW][64];
    const int n = threadIdx.x & 63, kb = threadIdx.x >> 6;
    float acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[j] = 0.f;
    float accb = 0.f;
    for (int r0 = blockIdx.x * RW; r0 < R; r0 += gridDim.x * RW) {
        __syncthreads();
        for (int i = threadIdx.x; i < RW * K; i += 256) {
            const int q = i / K, k = i % K, rr = r0 + q;
            float x = 0.f;
            if (rr < R) x = k < KA ? Act<T>::get(A + (size_t)rr * KA + k) : Act<T>::get(B + (size_t)rr * KB + (k - KA));
            cin[q][k] = x;
        }
        for (int i = threadIdx.x; i < RW * 64; i += 256) {
            const int q = i / 64, c = i % 64, rr = r0 + q;
            float gv = 0.f;
            if (rr < R) {
                gv = dOut[(size_t)rr * 64 + c];
                if (MASK && !(Act<T>::get(Hout + (size_t)rr * 64 + c) > 0.f)) gv = 0.f;
            }
            d[q][c] = gv;
        }
        __syncthreads();
        for (int q = 0; q < RW; ++q) {
            const float dq = d[q][n];
#pragma unroll
            for (int j = 0; j < NJ; ++j) acc[j] += cin[q][kb + 4 * j] * dq;
            accb += dq;
        }
    }
    float *o = part + (size_t)blockIdx.x * NOUT;
#pragma unroll
    for (int j = 0; j < NJ; ++j) o[(kb + 4 * j) * 64 + n] = acc[j];
    if (kb == 0) o[K * 64 + n] = accb;
}

__global__ void k_lp_wgrad_reduce(int nblk, int K, const float *__restrict__ part,
                                  float *__restrict__ gw, float *__restrict__ gb) {
    const int nout = K * 64 + 64;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * nout + o];
    if (o < K * 64) gw[o] = s;
    else gb[o - K * 64] = s;
}

__global__ void k_lp_adam(int n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 transposed copies of the four weight matrices.
__global__ void k_lp_wt(const float *__restrict__ w, uint16_t *__restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LP_NWT) return;
    int l = 3;
    while (l > 0 && i < LP_WT_OFF[l]) --l;
    const int K = l == 0 ? 32 : l == 3 ? 64 : 128;
    const int j = i - LP_WT_OFF[l], n = j / K, k = j % K;
    const int off = l == 3 ? LP_OFF_WH : LP_OFF_W[l];
    wt[i] = f32_to_bf16(w[off + k * 64 + n]);
}

// ------------------------------------------------------------ candidates

// Device-side control of one learned run.  The host enqueues iterations
// without reading anything back; the control kernels take the decisions of
// the iteration loop (stop, fall back to the exhaustive window, sizes), and
// every kernel of an iteration reads its counts and its skip flag here.
struct LPCtl {
    int done;       // the run has ended (converged, no donors, nothing to apply)
    int exh;        // this iteration scores every window entry exactly
    int skip_fwd;   // no forward pass this iteration
    int skip_norm;  // no sort of the one-per-donor candidates (done || exh)
    int skip_exh;   // no sort of the window candidates (done || !exh)
    int skip_swap;  // no swaps to apply (done || none accepted)
    int it;         // iterations run (the proposals' seeds follow it)
    int ndonor;     // donors counted by k_lp_pick
    int nd, nc, stride;
    int nacc, gain;  // accepted swaps and their summed gain (k_lp_select)
    int nsel;        // segments the accepted swaps touched
    int fallbacks, swaps;
    long long gain_total, exact_pairs;
};

// A run starts at iteration 0, in the exhaustive mode for mode 2, and is
// over at once when no segment is invalid.
__global__ void k_lp_ctl_init(LPCtl *ctl, const int32_t *__restrict__ tot, int mode) {
    if (threadIdx.x != 0) return;
    LPCtl z{};
    z.exh = mode == 2;
    z.done = tot[0] <= 0;
    *ctl = z;
}

__global__ void k_lp_ctl_begin(LPCtl *ctl, int mode) {
    if (threadIdx.x != 0) return;
    ctl->ndonor = 0;
    ctl->skip_fwd = ctl->done || ctl->exh || mode != 0;
}

// After the donors are counted: this iteration's sizes, or the end of the run
// when there is no donor.
__global__ void k_lp_ctl_mid(LPCtl *ctl, int topk, int W, int nseg) {
    if (threadIdx.x != 0) return;
    LPCtl c = *ctl;
    c.nd = min(topk, c.ndonor);
    if (!c.done && c.nd == 0) c.done = 1;
    if (c.done) {
        c.nc = 0;
        c.skip_norm = c.skip_exh = c.skip_swap = 1;
    } else {
        c.stride = max(1, nseg / c.nd);
        c.nc = c.exh ? c.nd * W : c.nd;
        c.exact_pairs += c.nc;
        c.skip_norm = c.exh;
        c.skip_exh = !c.exh;
    }
    *ctl = c;
}

// After the selection: nothing accepted falls back to the exhaustive window
// once (mode 0) or ends the run; otherwise the swaps are applied and the
// next iteration proposes again.
__global__ void k_lp_ctl_after(LPCtl *ctl, int mode, int32_t *__restrict__ queue) {
    if (threadIdx.x != 0) return;
    LPCtl c = *ctl;
    if (c.done) return;
    c.it++;
    if (c.nacc == 0) {
        c.skip_swap = 1;
        c.nsel = 0;
        if (mode == 0 && !c.exh) {
            c.exh = 1;
            c.fallbacks++;
        } else {
            c.done = 1;
        }
    } else {
        c.skip_swap = 0;
        c.nsel = 2 * c.nacc;
        c.swaps += c.nacc;
        c.gain_total += c.gain;
        if (mode == 0) c.exh = 0;
        *queue = 0;  // k_lp_seg_stats' work counter
    }
    *ctl = c;
}

// After the touched segments are recounted: over when nothing is invalid.
__global__ void k_lp_ctl_end(LPCtl *ctl, const int32_t *__restrict__ tot) {
    if (threadIdx.x == 0 && !ctl->done && !ctl->skip_swap && tot[0] <= 0) ctl->done = 1;
}

// The done flags of a batch's runs, in one array the host reads.
__global__ void k_lp_done_flags(LPCtl *const *__restrict__ ctls, int G, int32_t *__restrict__ done) {
    for (int i = threadIdx.x; i < G; i += blockDim.x) done[i] = ctls[i]->done;
}

// Per segment: the donor (a node of an invalid segment that sits in a row
// with N+1 ones, best donor score, ties to the lower bit) and the acceptor
// (best acceptor score).  mode 0: network scores Y; 1: seeded hash scores;
// 2: counts (donor = its rows with N+1 ones; acceptor = minus its degree and
// its segment's rows with at least N ones).  Keys sort ascending = best first.
__global__ void k_lp_pick(LPCtl *ctl, int nseg, int M, int mode0, uint64_t seed0, const int32_t *__restrict__ inv,
                          const int32_t *__restrict__ ptr, const int32_t *__restrict__ ccls,
                          const int32_t *__restrict__ segcls, const float *__restrict__ Y,
                          uint64_t *__restrict__ dkey, uint64_t *__restrict__ akey,
                          int8_t *__restrict__ dbest, int8_t *__restrict__ abest, int sbits) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || ctl->done) return;
    // the exhaustive window ranks by counts; the seed follows the iteration
    const int mode = ctl->exh ? 2 : mode0;
    const uint64_t seed = seed0 + (uint64_t)ctl->it * 0x9E3779B97F4A7C15ull;
    int32_t *ndonor = &ctl->ndonor;
    const int4 S = *(const int4 *)&segcls[s * 4];
    const bool invalid = S.z + S.w > 0;
    float bd = 0.f, ba = 0.f;
    int bdi = -1, bai = -1;
    for (int b = 0; b < M; ++b) {
        const int node = inv[s * M + b];
        float d, a;
        if (mode == 0) {
            d = Y[(size_t)node * LP_HD + 0];
            a = Y[(size_t)node * LP_HD + 1];
        } else if (mode == 1) {
            d = (float)(mix32(seed ^ ((uint64_t)node << 1)) >> 8);
            a = (float)(mix32(seed ^ ((uint64_t)node << 1) ^ 1ull) >> 8);
        } else {
            d = (float)ccls[node * 4 + 2];
            a = -(float)(ptr[node + 1] - ptr[node] + S.y + S.z + S.w);
        }
        if (invalid && ccls[node * 4 + 2] > 0 && (bdi < 0 || d > bd)) { bd = d; bdi = b; }
        if (bai < 0 || a > ba) { ba = a; bai = b; }
    }
    dbest[s] = (int8_t)bdi;
    abest[s] = (int8_t)bai;
    // (score descending, segment) in 32 + sbits bits, sbits the bits of the
    // largest segment index: the keys' order, sorted on those bits only
    // one array for one sort: donors first, acceptors after them (bit
    // 32 + sbits set), a segment without a donor the last of the donors
    const uint64_t low = (1ull << (32 + sbits)) - 1;
    dkey[s] = bdi >= 0 ? ((uint64_t)(~ord_f32(bd)) << sbits) | (uint32_t)s : low;
    akey[s] = (low + 1) | ((uint64_t)(~ord_f32(ba)) << sbits) | (uint32_t)s;
    // donors counted once per wave (lanes past nseg have returned)
    const uint64_t dm = __ballot(bdi >= 0);
    const uint64_t live = __ballot(1);
    if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)live) - 1 && dm)
        atomicAdd(ndonor, (int)__popcll(dm));
}

struct Cand { int p, q, bu, bv; };

// Donor i (in donor order) looks at the acceptors ranked (i*stride + w) mod
// nseg, w < W.  all == 0: one candidate per donor, the window entry with the
// best pair score (network: Y0[u] + Y1[v] + P[u].Q[v]; hash in mode 1; the
// first in mode 2); all == 1: every window entry is a candidate.
__global__ void k_lp_match(const LPCtl *__restrict__ ctl, int W, int nseg, int M, int mode0, uint64_t seed0,
                           const uint64_t *__restrict__ sdkey, const uint64_t *__restrict__ sakey,
                           const int8_t *__restrict__ dbest, const int8_t *__restrict__ abest,
                           const int32_t *__restrict__ inv, const float *__restrict__ Y,
                           Cand *__restrict__ cand, uint64_t smask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl->done || i >= ctl->nd) return;
    const int stride = ctl->stride, all = ctl->exh, mode = ctl->exh ? 2 : mode0;
    const uint64_t seed = seed0 ^ ((uint64_t)ctl->it << 20);
    const int p = (int)(sdkey[i] & smask);
    const int bu = dbest[p];
    const int u = inv[p * M + bu];
    float best = 0.f;
    int bw = -1;
    for (int w = 0; w < W; ++w) {
        const int j = (int)(((int64_t)i * stride + w) % nseg);
        const int q = (int)(sakey[j] & smask);
        const int bv = abest[q];
        Cand c{p, q == p ? -1 : q, bu, bv};
        if (all) {
            cand[(size_t)i * W + w] = c;
            continue;
        }
        if (q == p) continue;
        const int v = inv[q * M + bv];
        float sc;
        if (mode == 0) {
            const float *yu = Y + (size_t)u * LP_HD, *yv = Y + (size_t)v * LP_HD;
            sc = yu[0] + yv[1];
            for (int t = 0; t < LP_PQ; ++t) sc += yu[2 + t] * yv[2 + LP_PQ + t];
        } else if (mode == 1) {
            sc = (float)(mix32(seed ^ ((uint64_t)u << 32) ^ (uint64_t)v) >> 8);
        } else {
            sc = -(float)w;
        }
        if (bw < 0 || sc > best) { best = sc; bw = w; }
    }
    if (!all) {
        if (bw < 0) {
            cand[i] = Cand{p, -1, bu, 0};
        } else {
            const int j = (int)(((int64_t)i * stride + bw) % nseg);
            const int q = (int)(sakey[j] & smask);
            cand[i] = Cand{p, q, bu, abest[q]};
        }
    }
}

// ------------------------------------------------------------ exact pair scorer

// Exact M x M gains of candidate i's segment pair (p, q) from the entries of
// the two segments:
//   lgain[b1][b2] = E_p[b1] - Bl[b1][b2],  rgain[b1][b2] = E_q[b2] - Br[b1][b2]
//   E_s[b] = rows of s with popcount N+1 and bit b set
//   Bl = sum over rows in both of y_b2 * (x_b1 ? [pl = N+1] : [pl = N])
//   Br = sum over rows in both of x_b1 * (y_b2 ? [pr = N+1] : [pr = N])
// (the per-row closed form of GetSegSwapScore, reorder/reorder.h:108-122,
// summed; rows in neither segment or in one only contribute through E).
// Writes the best pair (max lgain + rgain, then max rgain, then the lower
// index b1*M+b2) and, when full != nullptr, the whole matrix in the layout
// of vnm_swap_score.
__global__ __launch_bounds__(256) void k_lp_pair_gain(const int *__restrict__ d_nc, int R, int M, int N,
                                                      const Cand *__restrict__ cand,
                                                      const int32_t *__restrict__ inv,
                                                      const int32_t *__restrict__ ptr,
                                                      const int32_t *__restrict__ idx,
                                                      int4 *__restrict__ best, int32_t *__restrict__ full) {
    __shared__ ColSpan cols[16];
    __shared__ int pre[17];
    __shared__ int sel[2];
    __shared__ int sBl[64], sBr[64], sEp[8], sEq[8];
    __shared__ int tlo[16], thi[16], tpre[17];
    __shared__ int tend[16][PB];
    // row bitmaps of the pair: an LDS hash (byte 0 = p, byte 1 = q) for at
    // most HASH_MAX entries, else LDS tiles of PAIR_TILE_ROWS rows, 16 bits
    // per row
    __shared__ union {
        RowHash h;
        uint32_t tile[PAIR_TILE_ROWS / 2];
    } L;
    const int i = blockIdx.x;
    if (d_nc && i >= *d_nc) return;  // past this iteration's candidates (device count)
    const Cand c = cand[i];
    const int tid = threadIdx.x, lane = tid & 63;
    const int MM = M * M;
    if (c.q < 0) {
        if (tid == 0) best[i] = make_int4(0, 0, -0x40000000, 0);
        return;
    }
    if (tid == 0) { sel[0] = c.p; sel[1] = c.q; }
    if (tid < 64) { sBl[tid] = 0; sBr[tid] = 0; }
    if (tid < 8) { sEp[tid] = 0; sEq[tid] = 0; }
    __syncthreads();
    const int total = table_columns(sel, 2, M, inv, ptr, cols, pre);
    const int b1 = lane / M, b2 = lane % M;
    int aBl = 0, aBr = 0, aEp = 0, aEq = 0;
    // one entry per lane (t = its column, -1 for none; lv / rv = its row's
    // bitmaps in p and q), folded into the gain sums with wave ballots; the
    // sums count each row once, through its owner (leading-bit) entries
    auto fold = [&](int t, uint32_t lv, uint32_t rv) {
        bool isP = false, isQ = false;
        uint32_t l = 0, rr = 0;
        if (t >= 0) {
            if (t < M) {
                l = lv;
                isP = (l >> (M - 1 - t)) == 1u;  // the row's owner entry in p
                if (isP) rr = rv;
            } else {
                rr = rv;
                isQ = (rr >> (M - 1 - (t - M))) == 1u;
            }
        }
        const int pl = popc32(l), pr = popc32(rr);
        const bool both = isP && rr != 0;
        const uint64_t Ap = __ballot(isP && pl == N + 1);
        const uint64_t Dq = __ballot(isQ && pr == N + 1);
        const uint64_t A = __ballot(both && pl == N + 1), B = __ballot(both && pl == N);
        const uint64_t C = __ballot(both && pr == N), D = __ballot(both && pr == N + 1);
        uint64_t Lb = 0, Rb = 0, Lown = 0, Rown = 0;
        for (int b = 0; b < M; ++b) {
            const uint64_t lm = __ballot(isP && ((l >> (M - 1 - b)) & 1u));
            const uint64_t rb = __ballot(both && ((rr >> (M - 1 - b)) & 1u));
            const uint64_t rq = __ballot(isQ && ((rr >> (M - 1 - b)) & 1u));
            if (b == b1) Lb = lm;
            if (b == b2) Rb = rb;
            if (b == lane) { Lown = lm; Rown = rq; }
        }
        if (lane < MM && (A | B | C | D)) {
            aBl += __popcll(Rb & ((Lb & A) | (~Lb & B)));
            aBr += __popcll(Lb & ((Rb & D) | (~Rb & C)));
        }
        if (lane < M) {
            aEp += __popcll(Lown & Ap);
            aEq += __popcll(Rown & Dq);
        }
    };
    if (total <= HASH_MAX) {
        hash_reset(L.h);
        __syncthreads();
        hash_scatter(L.h, M, cols, pre, 2 * M, total, idx);
        __syncthreads();
        // TAB_U entries per thread resolved before their lookups
        for (int base = 0; base < total; base += 256 * TAB_U) {
            int tt[TAB_U], rw[TAB_U];
            table_entries(cols, pre, 2 * M, total, idx, base + tid, 256, tt, rw);
#pragma unroll
            for (int u = 0; u < TAB_U; ++u) {
                const uint32_t h = tt[u] >= 0 ? hash_get(L.h, (uint32_t)rw[u]) : 0u;
                fold(tt[u], h & 0xFFu, (h >> 8) & 0xFFu);
            }
        }
    } else {
        // rows [r0, r0 + PAIR_TILE_ROWS) at a time: each column's entries of
        // a tile are one range of its row-sorted list; the ends of PB tiles
        // of all 2M columns come from one round of binary searches
        const int nc = 2 * M;
        for (int w = tid; w < PAIR_TILE_ROWS / 2; w += 256) L.tile[w] = 0u;
        if (tid < nc) tlo[tid] = cols[tid].start;
        const int ntile = (R + PAIR_TILE_ROWS - 1) / PAIR_TILE_ROWS;
        for (int g0 = 0; g0 < ntile; g0 += PB) {
            if (tid < nc * PB) {
                const int t = tid / PB, j = tid % PB;
                const int64_t r1 = min((int64_t)R, (int64_t)(g0 + j + 1) * PAIR_TILE_ROWS);
                int lo = cols[t].start, hi = cols[t].start + cols[t].len;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (idx[mid] < r1) lo = mid + 1;
                    else hi = mid;
                }
                tend[t][j] = lo;
            }
            __syncthreads();
            for (int j = 0; j < PB && g0 + j < ntile; ++j) {
                const int r0 = (g0 + j) * PAIR_TILE_ROWS;
                if (tid == 0) {
                    tpre[0] = 0;
                    for (int t = 0; t < nc; ++t) {
                        thi[t] = tend[t][j];
                        tpre[t + 1] = tpre[t] + thi[t] - tlo[t];
                    }
                }
                __syncthreads();
                const int tt = tpre[nc];
                if (tt > 0) {  // uniform
                    // TAB_U entries per thread resolved before their loads
                    auto entries = [&](int e0, int (&t)[TAB_U], int (&r)[TAB_U]) {
#pragma unroll
                        for (int u = 0; u < TAB_U; ++u) {
                            const int e = e0 + u * 256;
                            t[u] = e < tt ? entry_col(tpre, nc, e) : -1;
                        }
#pragma unroll
                        for (int u = 0; u < TAB_U; ++u)
                            r[u] = t[u] >= 0 ? idx[tlo[t[u]] + (e0 + u * 256 - tpre[t[u]])] - r0 : 0;
                    };
                    for (int e0 = tid; e0 < tt; e0 += 256 * TAB_U) {
                        int t[TAB_U], r[TAB_U];
                        entries(e0, t, r);
#pragma unroll
                        for (int u = 0; u < TAB_U; ++u)
                            if (t[u] >= 0)
                                atomicOr(&L.tile[r[u] >> 1],
                                         ((1u << (M - 1 - t[u] % M)) << (8 * (t[u] / M))) << (16 * (r[u] & 1)));
                    }
                    __syncthreads();
                    for (int base = 0; base < tt; base += 256 * TAB_U) {
                        int t[TAB_U], r[TAB_U];
                        entries(base + tid, t, r);
#pragma unroll
                        for (int u = 0; u < TAB_U; ++u) {
                            const uint32_t h = t[u] >= 0 ? (L.tile[r[u] >> 1] >> (16 * (r[u] & 1))) & 0xFFFFu : 0u;
                            fold(t[u], h & 0xFFu, h >> 8);
                        }
                    }
                    __syncthreads();
                    for (int w = tid; w < PAIR_TILE_ROWS / 2; w += 256) L.tile[w] = 0u;
                }
                if (tid < nc) tlo[tid] = thi[tid];
                __syncthreads();
            }
        }
    }
    if (lane < MM) {
        if (aBl) atomicAdd(&sBl[lane], aBl);
        if (aBr) atomicAdd(&sBr[lane], aBr);
    }
    if (lane < M) {
        if (aEp) atomicAdd(&sEp[lane], aEp);
        if (aEq) atomicAdd(&sEq[lane], aEq);
    }
    __syncthreads();
    if (tid < 64) {
        int lg = -0x40000000, rg = 0, j = tid;
        if (tid < MM) {
            lg = sEp[b1] - sBl[tid];
            rg = sEq[b2] - sBr[tid];
            if (full) {
                full[(size_t)i * 2 * MM + tid * 2 + 0] = lg;
                full[(size_t)i * 2 * MM + tid * 2 + 1] = rg;
            }
        }
        int sum = lg + rg;
        for (int off = 1; off < 64; off <<= 1) {
            const int os = __shfl_xor(sum, off), org = __shfl_xor(rg, off), oj = __shfl_xor(j, off);
            const int olg = __shfl_xor(lg, off);
            const bool take = os > sum || (os == sum && (org > rg || (org == rg && oj < j)));
            if (take) { sum = os; rg = org; j = oj; lg = olg; }
        }
        if (tid == 0) best[i] = make_int4(j / M, j % M, lg, rg);
    }
}

// ------------------------------------------------------------ selection

// (gain descending, candidate) in gbits + ibits bits: ibits the bits of the
// largest candidate index, gmax = 2^gbits - 1 at least the largest gain (a
// gain counts rows of the pair's two segments, each row at most once per
// segment, so it is at most twice the row count)
// Keys for the first nkeys slots; slots past this iteration's candidates get
// the sentinel ~0 (it sorts last), so the sort runs at a size fixed ahead.
__global__ void k_lp_gain_keys(const LPCtl *__restrict__ ctl, int nkeys, const Cand *__restrict__ cand,
                               const int4 *__restrict__ best, uint64_t *__restrict__ keys, int ibits,
                               uint32_t gmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nkeys || ctl->done) return;
    uint64_t key = ~0ull;
    if (i < ctl->nc) {
        const int4 b = best[i];
        const int gain = b.z + b.w;
        if (cand[i].q >= 0 && gain > 0) key = ((uint64_t)(gmax - (uint32_t)gain) << ibits) | (uint32_t)i;
    }
    keys[i] = key;
}

constexpr int SEL_T = 1024, SEL_B = 1024;
constexpr int SEL_ROUNDS = 24;                 // parallel rounds before the sequential finish
constexpr int32_t OWNER_NONE = 0x7F7F7F7F;    // empty owner (bytes 0x7F, set by a memset)

// Accept candidates in (gain desc, index asc) order when both segments are
// still unused: the sequential greedy (reorder/reorder.h's segment-disjoint
// rule applied to the learned batch), computed in rounds.  In a round every
// live candidate k writes min(k) into the owner slot of its two segments;
// k is accepted when it owns both (no live candidate before it shares a
// segment, so the sequential walk would accept it too), and candidates
// touching a segment used by now are dropped.  The first live candidate is
// always accepted, so rounds make progress; after SEL_ROUNDS rounds thread 0
// finishes the chunk sequentially.  Chunks of SEL_B candidates in order.
// out[0] = accepted, out[1] = their total gain; acc = candidate indices in
// order.  owner: nseg ints, OWNER_NONE on entry and on exit.
__global__ __launch_bounds__(SEL_T) void k_lp_select(LPCtl *ctl, int nseg, const uint64_t *__restrict__ skeys,
                                                     const Cand *__restrict__ cand,
                                                     const int4 *__restrict__ best, int32_t *__restrict__ acc,
                                                     int32_t *owner, uint64_t imask) {
    if (ctl->done) return;
    const int nc = ctl->nc;
    extern __shared__ uint32_t used[];
    __shared__ int bp[SEL_B], bq[SEL_B], bg[SEL_B], bc[SEL_B];
    __shared__ uint8_t live[SEL_B], took[SEL_B];
    __shared__ int wsum[SEL_T / 64], wg[SEL_T / 64];
    __shared__ int nacc, gsum, stop;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int words = (nseg + 31) / 32;
    for (int i = tid; i < words; i += SEL_T) used[i] = 0;
    if (tid == 0) { nacc = 0; gsum = 0; stop = 0; }
    __syncthreads();
    for (int base = 0; base < nc && !stop; base += SEL_B) {
        const int n = min(SEL_B, nc - base);
        // one candidate per thread (SEL_T == SEL_B)
        const int k = tid;
        bool alive = false;
        int p = 0, q = 0;
        if (k < n) {
            const uint64_t key = skeys[base + k];
            if (key != ~0ull) {
                const int ci = (int)(key & imask);
                p = cand[ci].p;
                q = cand[ci].q;
                const int4 b = best[ci];
                bp[k] = p;
                bq[k] = q;
                bg[k] = b.z + b.w;
                bc[k] = ci;
                alive = !((used[p >> 5] >> (p & 31)) & 1u) && !((used[q >> 5] >> (q & 31)) & 1u);
            } else {
                stop = 1;  // sorted keys: the rest are not candidates
            }
        }
        bool taken = false;
        for (int r = 0; r < SEL_ROUNDS; ++r) {
            if (!__syncthreads_or(alive)) break;
            if (alive) {
                atomicMin(&owner[p], k);
                atomicMin(&owner[q], k);
            }
            __threadfence();
            __syncthreads();
            bool win = false;
            if (alive) win = ld_l2(&owner[p]) == k && ld_l2(&owner[q]) == k;
            __syncthreads();
            if (alive) {
                st_l2(&owner[p], OWNER_NONE);
                st_l2(&owner[q], OWNER_NONE);
                if (win) {
                    taken = true;
                    atomicOr(&used[p >> 5], 1u << (p & 31));
                    atomicOr(&used[q >> 5], 1u << (q & 31));
                }
            }
            __threadfence();
            __syncthreads();
            if (alive) alive = !taken && !((used[p >> 5] >> (p & 31)) & 1u) && !((used[q >> 5] >> (q & 31)) & 1u);
        }
        if (k < SEL_B) {
            live[k] = alive;
            took[k] = taken;
        }
        // candidates left after the rounds: the sequential walk, in order
        if (__syncthreads_or(alive) && tid == 0) {
            for (int j = 0; j < n; ++j) {
                if (!live[j]) continue;
                const int a = bp[j], b = bq[j];
                if (((used[a >> 5] >> (a & 31)) & 1u) || ((used[b >> 5] >> (b & 31)) & 1u)) continue;
                used[a >> 5] |= 1u << (a & 31);
                used[b >> 5] |= 1u << (b & 31);
                took[j] = 1;
            }
        }
        __syncthreads();
        // accepted candidates in order: wave ballots, then wave offsets
        const bool t = k < n && took[k];
        const uint64_t bal = __ballot(t);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        int g = t ? bg[k] : 0;
        for (int d = 32; d >= 1; d >>= 1) g += __shfl_xor(g, d, 64);
        if (lane == 0) {
            wsum[w] = __popcll(bal);
            wg[w] = g;
        }
        __syncthreads();
        int off = nacc;
        for (int v = 0; v < w; ++v) off += wsum[v];
        if (t) acc[off + before] = bc[k];
        __syncthreads();
        if (tid == 0) {
            for (int v = 0; v < SEL_T / 64; ++v) {
                nacc += wsum[v];
                gsum += wg[v];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        ctl->nacc = nacc;
        ctl->gain = gsum;
    }
}

// Apply the accepted swaps (segment-disjoint, so in parallel) and list the
// touched segments.
__global__ void k_lp_swap(const LPCtl *__restrict__ ctl, int M, const int32_t *__restrict__ acc,
                          const Cand *__restrict__ cand, const int4 *__restrict__ best, int32_t *__restrict__ perm,
                          int32_t *__restrict__ inv, int32_t *__restrict__ touched) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl->skip_swap || t >= ctl->nacc) return;
    const int ci = acc[t];
    const Cand c = cand[ci];
    const int4 b = best[ci];
    const int xa = c.p * M + b.x, xb = c.q * M + b.y;
    const int na = inv[xa], nb = inv[xb];
    inv[xa] = nb;
    inv[xb] = na;
    perm[na] = xb;
    perm[nb] = xa;
    touched[2 * t] = c.p;
    touched[2 * t + 1] = c.q;
}

// ------------------------------------------------------------ host: cache

struct AggPlan {
    int nheavy = 0, nch = 0;
    DBuf<int32_t> heavy, ch0, beg, end;
    DBuf<float> part;  // nch x 64
};

struct LPCache {
    int64_t epoch = -1;
    int R = 0, nseg = 0, M = 0;
    int64_t nadj = 0;
    DBuf<int32_t> ptr, idx, tptr, tidx;
    bool has_bwd = false;  // tptr/tidx and the backward plan are built
    DBuf<float> inv_deg;
    AggPlan fwd, bwd;
    DBuf<int32_t> ccls, segcls, segnnz, tot;
    DBuf<float> X, A0, H1, A1, H2, A2, H3, Y;  // sized for FP32, BF16 uses the front half
    DBuf<float> dY, dH, dIn, dAg, part, lpart;
    DBuf<int32_t> pairs;
    DBuf<float> target;
};

static void lp_cache_free(void *p) { delete (LPCache *)p; }

static int agg_plan(const std::vector<int32_t> &hptr, AggPlan &P) {
    std::vector<int32_t> heavy, ch0{0}, beg, end;
    const int R = (int)hptr.size() - 1;
    for (int u = 0; u < R; ++u) {
        const int b = hptr[u], e = hptr[u + 1];
        if (e - b <= AGG_CH) continue;
        heavy.push_back(u);
        for (int s = b; s < e; s += AGG_CH) {
            beg.push_back(s);
            end.push_back(std::min(e, s + AGG_CH));
        }
        ch0.push_back((int)beg.size());
    }
    P.nheavy = (int)heavy.size();
    P.nch = (int)beg.size();
    int rc;
    if ((rc = P.heavy.alloc(std::max(1, P.nheavy))) || (rc = P.ch0.alloc(P.nheavy + 1)) ||
        (rc = P.beg.alloc(std::max(1, P.nch))) || (rc = P.end.alloc(std::max(1, P.nch))) ||
        (rc = P.part.alloc((size_t)std::max(1, P.nch) * 64)))
        return rc;
    if (P.nheavy) {
        VNM_HIP_TRY(hipMemcpy(P.heavy.p, heavy.data(), heavy.size() * 4, hipMemcpyHostToDevice));
        VNM_HIP_TRY(hipMemcpy(P.beg.p, beg.data(), beg.size() * 4, hipMemcpyHostToDevice));
        VNM_HIP_TRY(hipMemcpy(P.end.p, end.data(), end.size() * 4, hipMemcpyHostToDevice));
    }
    VNM_HIP_TRY(hipMemcpy(P.ch0.p, ch0.data(), ch0.size() * 4, hipMemcpyHostToDevice));
    return VNM_OK;
}

static int build_lists(vnm_graph *g, uint64_t *keys, uint64_t *tmp, int64_t n, int64_t nvalid,
                       DBuf<int32_t> &ptr, DBuf<int32_t> &idx, std::vector<int32_t> &hptr) {
    hipStream_t st = g->ctx->stream;
    const int R = g->nrows;
    int bits = 1;
    while ((1ll << bits) <= R) ++bits;
    int rc = sort_keys_u64(keys, tmp, n, 32 + bits, st);
    if (rc) return rc;
    if ((rc = ptr.alloc(R + 1)) || (rc = idx.alloc(std::max<int64_t>(nvalid, 1)))) return rc;
    VNM_HIP_TRY(hipMemsetAsync(ptr.p, 0, (R + 1) * 4, st));  // stays all zero when nothing is valid
    if (nvalid > 0)
        hipLaunchKernelGGL(k_lp_adj_unpack, dim3(grid_of(nvalid, 256)), dim3(256), 0, st, tmp, nvalid, R,
                           idx.p, ptr.p);
    g->ctx->launches += 1;
    VNM_LAUNCH_CHECK("k_lp_adj_unpack");
    hptr.resize(R + 1);
    VNM_HIP_TRY(hipMemcpyAsync(hptr.data(), ptr.p, (R + 1) * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

// The proposer's view of the graph at its current labelling: sorted
// adjacency both ways, aggregation plans, statistics and activation buffers.
static int lp_prepare(vnm_graph *g, LPCache **out) {
    LPCache *c = (LPCache *)g->lp_cache;
    if (c && c->epoch == g->epoch && c->R == g->nrows && c->M == g->M) {
        *out = c;
        return VNM_OK;
    }
    if (c) { delete c; g->lp_cache = nullptr; }
    c = new LPCache();
    g->lp_cache = c;
    g->lp_cache_free = lp_cache_free;
    hipStream_t st = g->ctx->stream;
    const int R = g->nrows;
    c->R = R;
    c->M = g->M;
    c->nseg = g->nseg();
    const int64_t nnz = g->nnz;
    DBuf<uint64_t> kc, kt, tmp;
    DBuf<int32_t> nv;
    int rc;
    if ((rc = kc.alloc(std::max<int64_t>(nnz, 1))) || (rc = kt.alloc(std::max<int64_t>(nnz, 1))) ||
        (rc = tmp.alloc(std::max<int64_t>(nnz, 1))) || (rc = nv.alloc(1)))
        return rc;
    VNM_HIP_TRY(hipMemsetAsync(nv.p, 0, 4, st));
    if (nnz)
        hipLaunchKernelGGL(k_lp_adj_keys, dim3(grid_of(nnz, 256, 4096)), dim3(256), 0, st, g->row.p,
                           g->col.p, g->val.p, nnz, R, kc.p, kt.p, nv.p);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_adj_keys");
    int32_t nvalid = 0;
    VNM_HIP_TRY(hipMemcpyAsync(&nvalid, nv.p, 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    c->nadj = nvalid;
    std::vector<int32_t> hp;
    if ((rc = build_lists(g, kc.p, tmp.p, nnz, nvalid, c->ptr, c->idx, hp))) return rc;
    if ((rc = c->inv_deg.alloc(R))) return rc;
    hipLaunchKernelGGL(k_lp_inv_deg, dim3(grid_of(R, 256)), dim3(256), 0, st, R, c->ptr.p, c->inv_deg.p);
    if ((rc = agg_plan(hp, c->fwd))) return rc;
    const int ns = c->nseg;
    if ((rc = c->ccls.alloc((size_t)R * 4)) || (rc = c->segcls.alloc((size_t)ns * 4)) ||
        (rc = c->segnnz.alloc(ns)) || (rc = c->tot.alloc(4)))
        return rc;
    const size_t R64 = (size_t)R * 64, R16 = (size_t)R * 16;
    if ((rc = c->X.alloc(R16)) || (rc = c->A0.alloc(R16)) || (rc = c->H1.alloc(R64)) ||
        (rc = c->A1.alloc(R64)) || (rc = c->H2.alloc(R64)) || (rc = c->A2.alloc(R64)) ||
        (rc = c->H3.alloc(R64)) || (rc = c->Y.alloc(R64)))
        return rc;
    c->epoch = g->epoch;
    *out = c;
    return VNM_OK;
}

// Class statistics over all segments (segs == nullptr) or a list.
static int lp_stats(vnm_graph *g, LPCache *c, int N, const int32_t *segs, int nsel) {
    const int n = segs ? nsel : c->nseg;
    if (n == 0) return VNM_OK;
    const int grid = std::max(1, std::min(n, 1024));
    hipStream_t st = g->ctx->stream;
    int rc;
    if (!g->queue.p && (rc = g->queue.alloc(4))) return rc;
    VNM_HIP_TRY(hipMemsetAsync(g->queue.p, 0, 4, st));
    hipLaunchKernelGGL(k_lp_seg_stats, dim3(grid), dim3(SS_T), 0, st, g->nrows, g->M, N, c->nseg, segs, nsel,
                       g->inv.p, c->ptr.p, c->idx.p, c->ccls.p, c->segcls.p, c->segnnz.p, g->queue.p, nullptr,
                       nullptr);
    hipLaunchKernelGGL(k_lp_totals, dim3(1), dim3(1024), 0, st, c->nseg, c->segcls.p, c->tot.p, nullptr);
    g->ctx->launches += 2;
    VNM_LAUNCH_CHECK("k_lp_seg_stats");
    return VNM_OK;
}

// The transposed adjacency and its aggregation plan, which only backward
// passes read: built on first use (once per labelling, like the rest).
static int lp_prepare_bwd(vnm_graph *g, LPCache *c) {
    if (c->has_bwd) return VNM_OK;
    hipStream_t st = g->ctx->stream;
    const int64_t nnz = g->nnz;
    DBuf<uint64_t> kc, kt, tmp;
    DBuf<int32_t> nv;
    int rc;
    if ((rc = kc.alloc(std::max<int64_t>(nnz, 1))) || (rc = kt.alloc(std::max<int64_t>(nnz, 1))) ||
        (rc = tmp.alloc(std::max<int64_t>(nnz, 1))) || (rc = nv.alloc(1)))
        return rc;
    VNM_HIP_TRY(hipMemsetAsync(nv.p, 0, 4, st));
    if (nnz)
        hipLaunchKernelGGL(k_lp_adj_keys, dim3(grid_of(nnz, 256, 4096)), dim3(256), 0, st, g->row.p,
                           g->col.p, g->val.p, nnz, c->R, kc.p, kt.p, nv.p);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_adj_keys");
    std::vector<int32_t> htp;
    if ((rc = build_lists(g, kt.p, tmp.p, nnz, c->nadj, c->tptr, c->tidx, htp))) return rc;
    if ((rc = agg_plan(htp, c->bwd))) return rc;
    c->has_bwd = true;
    return VNM_OK;
}

template <typename TI, typename TO, int F, bool BWD>
static int lp_aggregate(vnm_graph *g, LPCache *c, const TI *H, TO *out, const int *skip = nullptr) {
    hipStream_t st = g->ctx->stream;
    const int R = c->R;
    if (BWD) {
        const int rc = lp_prepare_bwd(g, c);
        if (rc) return rc;
    }
    const AggPlan &P = BWD ? c->bwd : c->fwd;
    const int32_t *ptr = BWD ? c->tptr.p : c->ptr.p;
    const int32_t *idx = BWD ? c->tidx.p : c->idx.p;
    hipLaunchKernelGGL((k_lp_agg_nodes<TI, TO, F, BWD>), dim3((R + 3) / 4), dim3(256), 0, st, R, ptr, idx,
                       c->inv_deg.p, H, out, skip);
    g->ctx->launches++;
    if (P.nch) {
        hipLaunchKernelGGL((k_lp_agg_chunks<TI, F, BWD>), dim3((P.nch + 3) / 4), dim3(256), 0, st, P.nch,
                           P.beg.p, P.end.p, idx, c->inv_deg.p, H, P.part.p, skip);
        hipLaunchKernelGGL((k_lp_agg_combine<TO, F, BWD>), dim3(grid_of((int64_t)P.nheavy * F, 256)),
                           dim3(256), 0, st, P.nheavy, P.heavy.p, P.ch0.p, P.part.p, c->inv_deg.p, out, skip);
        g->ctx->launches += 2;
    }
    VNM_LAUNCH_CHECK("k_lp_agg");
    return VNM_OK;
}

// Features, three layers and the head at the current perm/inv; Y is FP32.
// skip: a device flag; while it is nonzero every kernel returns at once.
static int lp_forward(vnm_graph *g, LPCache *c, const vnm_lp_model *m, int bf16, const int *skip = nullptr) {
    hipStream_t st = g->ctx->stream;
    const int R = c->R;
    int rc;
    KernelTimer kt(g->ctx, "lp_forward");
    if (bf16) {
        typedef uint16_t T;
        T *X = (T *)c->X.p, *A0 = (T *)c->A0.p, *H1 = (T *)c->H1.p, *A1 = (T *)c->A1.p;
        T *H2 = (T *)c->H2.p, *A2 = (T *)c->A2.p, *H3 = (T *)c->H3.p;
        hipLaunchKernelGGL(k_lp_features<T>, dim3(grid_of(R, 256)), dim3(256), 0, st, R, g->M, g->perm.p,
                           c->ptr.p, c->ccls.p, c->segcls.p, c->segnnz.p, c->tot.p, X, skip);
        if ((rc = lp_aggregate<T, T, 16, false>(g, c, X, A0, skip))) return rc;
        const int gr = (R + 63) / 64;
        hipLaunchKernelGGL((k_lp_dense_mfma<16, 16, true, T>), dim3(gr), dim3(256), 0, st, R, X, A0,
                           m->wt.p + LP_WT_OFF[0], m->w.p + LP_OFF_B[0], H1, skip);
        if ((rc = lp_aggregate<T, T, 64, false>(g, c, H1, A1, skip))) return rc;
        hipLaunchKernelGGL((k_lp_dense_mfma<64, 64, true, T>), dim3(gr), dim3(256), 0, st, R, H1, A1,
                           m->wt.p + LP_WT_OFF[1], m->w.p + LP_OFF_B[1], H2, skip);
        if ((rc = lp_aggregate<T, T, 64, false>(g, c, H2, A2, skip))) return rc;
        hipLaunchKernelGGL((k_lp_dense_mfma<64, 64, true, T>), dim3(gr), dim3(256), 0, st, R, H2, A2,
                           m->wt.p + LP_WT_OFF[2], m->w.p + LP_OFF_B[2], H3, skip);
        hipLaunchKernelGGL((k_lp_dense_mfma<64, 0, false, float>), dim3(gr), dim3(256), 0, st, R, H3,
                           (const T *)nullptr, m->wt.p + LP_WT_OFF[3], m->w.p + LP_OFF_BH, c->Y.p, skip);
    } else {
        float *X = c->X.p, *A0 = c->A0.p, *H1 = c->H1.p, *A1 = c->A1.p, *H2 = c->H2.p, *A2 = c->A2.p,
              *H3 = c->H3.p;
        hipLaunchKernelGGL(k_lp_features<float>, dim3(grid_of(R, 256)), dim3(256), 0, st, R, g->M,
                           g->perm.p, c->ptr.p, c->ccls.p, c->segcls.p, c->segnnz.p, c->tot.p, X, skip);
        const int gr = grid_of(R, 4, 4096);
        if ((rc = lp_aggregate<float, float, 16, false>(g, c, X, A0, skip))) return rc;
        hipLaunchKernelGGL((k_lp_dense_f32<16, 16, true>), dim3(gr), dim3(256), 0, st, R, X, A0,
                           m->w.p + LP_OFF_W[0], m->w.p + LP_OFF_B[0], H1, skip);
        if ((rc = lp_aggregate<float, float, 64, false>(g, c, H1, A1, skip))) return rc;
        hipLaunchKernelGGL((k_lp_dense_f32<64, 64, true>), dim3(gr), dim3(256), 0, st, R, H1, A1,
                           m->w.p + LP_OFF_W[1], m->w.p + LP_OFF_B[1], H2, skip);
        if ((rc = lp_aggregate<float, float, 64, false>(g, c, H2, A2, skip))) return rc;
        hipLaunchKernelGGL((k_lp_dense_f32<64, 64, true>), dim3(gr), dim3(256), 0, st, R, H2, A2,
                           m->w.p + LP_OFF_W[2], m->w.p + LP_OFF_B[2], H3, skip);
        hipLaunchKernelGGL((k_lp_dense_f32<64, 0, false>), dim3(gr), dim3(256), 0, st, R, H3,
                           (const float *)nullptr, m->w.p + LP_OFF_WH, m->w.p + LP_OFF_BH, c->Y.p, skip);
    }
    kt.stop();
    g->ctx->launches += 6;
    VNM_LAUNCH_CHECK("lp_forward");
    return VNM_OK;
}

template <typename T>
static int lp_backward_t(vnm_graph *g, LPCache *c, vnm_lp_model *m) {
    hipStream_t st = g->ctx->stream;
    const int R = c->R;
    const T *X = (const T *)c->X.p, *A0 = (const T *)c->A0.p, *H1 = (const T *)c->H1.p;
    const T *A1 = (const T *)c->A1.p, *H2 = (const T *)c->H2.p, *A2 = (const T *)c->A2.p;
    const T *H3 = (const T *)c->H3.p;
    float *G = m->g.p;
    const int gb = std::min(WG_GRID, std::max(1, (R + 15) / 16));
    const int gi = grid_of(R, 4, 4096);
    int rc;
    // head: dY -> dWh, dbh, dH3
    hipLaunchKernelGGL((k_lp_wgrad<T, 64, 0, false>), dim3(gb), dim3(256), 0, st, R, H3, (const T *)nullptr,
                       c->dY.p, (const T *)nullptr, c->part.p);
    hipLaunchKernelGGL(k_lp_wgrad_reduce, dim3(grid_of(64 * 64 + 64, 256)), dim3(256), 0, st, gb, 64,
                       c->part.p, G + LP_OFF_WH, G + LP_OFF_BH);
    hipLaunchKernelGGL((k_lp_bwd_in<T, 64, 0, false>), dim3(gi), dim3(256), 0, st, R, c->dY.p,
                       (const T *)nullptr, m->w.p + LP_OFF_WH, c->dH.p, (float *)nullptr);
    // layer 2: dH = dH3
    hipLaunchKernelGGL((k_lp_wgrad<T, 64, 64, true>), dim3(gb), dim3(256), 0, st, R, H2, A2, c->dH.p, H3,
                       c->part.p);
    hipLaunchKernelGGL(k_lp_wgrad_reduce, dim3(grid_of(128 * 64 + 64, 256)), dim3(256), 0, st, gb, 128,
                       c->part.p, G + LP_OFF_W[2], G + LP_OFF_B[2]);
    hipLaunchKernelGGL((k_lp_bwd_in<T, 64, 64, true>), dim3(gi), dim3(256), 0, st, R, c->dH.p, H3,
                       m->w.p + LP_OFF_W[2], c->dIn.p, c->dAg.p);
    if ((rc = lp_aggregate<float, float, 64, true>(g, c, c->dAg.p, c->dIn.p))) return rc;
    // layer 1: dH2 in dIn
    hipLaunchKernelGGL((k_lp_wgrad<T, 64, 64, true>), dim3(gb), dim3(256), 0, st, R, H1, A1, c->dIn.p, H2,
                       c->part.p);
    hipLaunchKernelGGL(k_lp_wgrad_reduce, dim3(grid_of(128 * 64 + 64, 256)), dim3(256), 0, st, gb, 128,
                       c->part.p, G + LP_OFF_W[1], G + LP_OFF_B[1]);
    hipLaunchKernelGGL((k_lp_bwd_in<T, 64, 64, true>), dim3(gi), dim3(256), 0, st, R, c->dIn.p, H2,
                       m->w.p + LP_OFF_W[1], c->dH.p, c->dAg.p);
    if ((rc = lp_aggregate<float, float, 64, true>(g, c, c->dAg.p, c->dH.p))) return rc;
    // layer 0: dH1 in dH (input features need no gradient)
    hipLaunchKernelGGL((k_lp_wgrad<T, 16, 16, true>), dim3(gb), dim3(256), 0, st, R, X, A0, c->dH.p, H1,
                       c->part.p);
    hipLaunchKernelGGL(k_lp_wgrad_reduce, dim3(grid_of(32 * 64 + 64, 256)), dim3(256), 0, st, gb, 32,
                       c->part.p, G + LP_OFF_W[0], G + LP_OFF_B[0]);
    g->ctx->launches += 13;
    VNM_LAUNCH_CHECK("lp_backward");
    return VNM_OK;
}

// Loss of the pair head on (pairs, target) and, when grad, the full gradient
// into m->g.  Pairs are current ids.
static int lp_loss_grad(vnm_graph *g, LPCache *c, vnm_lp_model *m, int bf16, const int32_t *pairs,
                        const float *target, int count, bool grad, float *loss) {
    hipStream_t st = g->ctx->stream;
    const int R = c->R;
    int rc;
    if (c->pairs.n < (size_t)std::max(2 * count, 2)) {
        c->pairs.release();
        c->target.release();
        if ((rc = c->pairs.alloc(std::max(2 * count, 2))) || (rc = c->target.alloc(std::max(count, 1))))
            return rc;
    }
    if (!c->dY.p) {
        const size_t R64 = (size_t)R * 64;
        if ((rc = c->dY.alloc(R64)) || (rc = c->dH.alloc(R64)) || (rc = c->dIn.alloc(R64)) ||
            (rc = c->dAg.alloc(R64)) || (rc = c->part.alloc((size_t)WG_GRID * (128 * 64 + 64))) ||
            (rc = c->lpart.alloc(65536)))
            return rc;
    }
    const int nb = (count + 255) / 256;
    if (nb > 65536) return set_error(VNM_E_ARG, "too many training pairs (%d)", count);
    if (count) {
        VNM_HIP_TRY(hipMemcpyAsync(c->pairs.p, pairs, (size_t)count * 8, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemcpyAsync(c->target.p, target, (size_t)count * 4, hipMemcpyHostToDevice, st));
    }
    VNM_HIP_TRY(hipMemsetAsync(c->dY.p, 0, (size_t)R * 64 * 4, st));
    if (count)
        hipLaunchKernelGGL(k_lp_pair_loss, dim3(nb), dim3(256), 0, st, count, c->pairs.p, c->target.p, c->Y.p,
                           c->dY.p, c->lpart.p);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_pair_loss");
    if (grad) {
        rc = bf16 ? lp_backward_t<uint16_t>(g, c, m) : lp_backward_t<float>(g, c, m);
        if (rc) return rc;
    }
    std::vector<float> lp(std::max(nb, 1));
    if (count) VNM_HIP_TRY(hipMemcpyAsync(lp.data(), c->lpart.p, nb * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    double s = 0;
    for (int b = 0; b < nb; ++b) s += lp[b];
    if (loss) *loss = count ? (float)(s / count) : 0.f;
    return VNM_OK;
}

static int lp_check(vnm_graph *g, int N) {
    if (g->M != 4 && g->M != 8) return set_error(VNM_E_ARG, "the proposer supports M = 4 or 8 (got %d)", g->M);
    if (N < 1 || N >= g->M) return set_error(VNM_E_ARG, "N must be in [1, M) (got %d)", N);
    if (g->nrows % g->M != 0 || g->ncols > g->nrows)
        return set_error(VNM_E_ARG, "the proposer needs nrows %% M == 0 and ncols <= nrows");
    return VNM_OK;
}

// A labelling-fresh start: identity perm/inv over the prepared adjacency and
// statistics of every segment.
static int lp_begin(vnm_graph *g, int N, LPCache **out) {
    int rc = lp_check(g, N);
    if (rc) return rc;
    if ((rc = graph_scratch(g))) return rc;
    LPCache *c;
    if ((rc = lp_prepare(g, &c))) return rc;
    hipLaunchKernelGGL(k_lp_iota2, dim3(grid_of(g->nrows, 256)), dim3(256), 0, g->ctx->stream, g->nrows,
                       g->perm.p, g->inv.p);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_iota2");
    if ((rc = lp_stats(g, c, N, nullptr, 0))) return rc;
    *out = c;
    return VNM_OK;
}

// ------------------------------------------------------------ host: reorder

struct LPRound {
    DBuf<uint64_t> dkey, sdkey, gkey, sgkey;  // dkey/sdkey: donors then acceptors (2 x nseg)
    DBuf<int8_t> dbest, abest;
    DBuf<Cand> cand;
    DBuf<int4> best;
    DBuf<int32_t> acc, out, touched, owner;
};

// Score candidates [0, nc) exactly: a workgroup per candidate.
// nc candidates, or (d_nc given) the first *d_nc of nc launched blocks.
static int lp_score_cands(vnm_graph *g, LPCache *c, int N, Cand *cand, int4 *best, int nc, int32_t *full,
                          const int *d_nc = nullptr) {
    if (nc <= 0) return VNM_OK;
    KernelTimer kt(g->ctx, "k_lp_pair_gain");
    hipLaunchKernelGGL(k_lp_pair_gain, dim3(nc), dim3(256), 0, g->ctx->stream, d_nc, g->nrows, g->M, N, cand,
                       g->inv.p, c->ptr.p, c->idx.p, best, full);
    kt.stop();
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_pair_gain");
    return VNM_OK;
}

static int lp_reorder_args(vnm_graph *g, const vnm_lp_model *m, int mode, int topk, int window, int maxiter) {
    if (mode < 0 || mode > 2) return set_error(VNM_E_ARG, "mode must be 0 (learned), 1 (random) or 2 (exhaustive)");
    if (mode == 0 && !m) return set_error(VNM_E_ARG, "learned mode needs a model");
    if (m && m->ctx != g->ctx) return set_error(VNM_E_ARG, "model and graph belong to different contexts");
    if (topk < 1 || window < 1 || maxiter < 0) return set_error(VNM_E_ARG, "topk, window >= 1, maxiter >= 0");
    if (g->nseg() > (1 << 18)) return set_error(VNM_E_ARG, "more than 2^18 segments");
    return VNM_OK;
}

// One learned run, split so that its iterations can be enqueued without the
// host reading anything back: begin (cache, buffers, control block), any
// number of iterations, end (statistics, relabel).
struct LPRun {
    vnm_graph *g = nullptr;
    const vnm_lp_model *m = nullptr;
    LPCache *c = nullptr;
    int N = 0, mode = 0, topk = 0, window = 0, ns = 0, M = 0;
    int nnorm = 0;  // keys sorted in an iteration with one candidate per donor
    int ncap = 0;   // ... and with the whole window (the most candidates)
    int sbits = 0, ibits = 0, gbits = 0;
    int init_cnt = 0, init_size = 0;
    uint32_t gmax = 0;
    uint64_t smask = 0, imask = 0, seed = 0;
    LPRound w;
    DBuf<LPCtl> ctl;
    std::chrono::steady_clock::time_point t0;
};

static int lp_run_begin(LPRun &r, vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window,
                        uint64_t seed) {
    r.t0 = std::chrono::steady_clock::now();
    r.g = g;
    r.m = m;
    r.N = N;
    r.mode = mode;
    r.topk = topk;
    r.window = window;
    r.seed = seed;
    int rc = lp_begin(g, N, &r.c);
    if (rc) return rc;
    hipStream_t st = g->ctx->stream;
    const int ns = r.ns = r.c->nseg;
    r.M = g->M;
    r.nnorm = std::max(1, std::min(topk, ns));
    r.ncap = mode == 1 ? r.nnorm : r.nnorm * window;
    // sort keys hold a segment or candidate index in their low bits
    auto bits_of = [](int64_t x) { int b = 1; while ((1ll << b) <= x) ++b; return b; };
    r.sbits = bits_of(ns - 1);
    r.ibits = bits_of(r.ncap - 1);
    r.gbits = bits_of(2 * (int64_t)g->nrows);
    r.gmax = (uint32_t)((1ull << r.gbits) - 1);
    r.smask = (1ull << r.sbits) - 1;
    r.imask = (1ull << r.ibits) - 1;
    LPRound &w = r.w;
    if ((rc = w.dkey.alloc(2 * (size_t)ns)) || (rc = w.sdkey.alloc(2 * (size_t)ns)) ||
        (rc = w.dbest.alloc(ns)) || (rc = w.abest.alloc(ns)) ||
        (rc = w.cand.alloc(r.ncap)) || (rc = w.best.alloc(r.ncap)) || (rc = w.gkey.alloc(r.ncap)) ||
        (rc = w.sgkey.alloc(r.ncap)) || (rc = w.acc.alloc(r.ncap)) ||
        (rc = w.touched.alloc(2 * (size_t)std::min(r.ncap, ns))) || (rc = w.owner.alloc(ns)) ||
        (rc = r.ctl.alloc(1)))
        return rc;
    if (!g->queue.p && (rc = g->queue.alloc(4))) return rc;
    // the sorts of an iteration allocate nothing (they may run under capture)
    if ((rc = sort_scratch_reserve(std::max(2 * ns, r.ncap)))) return rc;
    VNM_HIP_TRY(hipMemsetAsync(w.owner.p, 0x7F, (size_t)ns * 4, st));  // OWNER_NONE in every slot
    int32_t tot[2];
    VNM_HIP_TRY(hipMemcpyAsync(tot, r.c->tot.p, 8, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    r.init_cnt = tot[0];
    r.init_size = tot[1];
    hipLaunchKernelGGL(k_lp_ctl_init, dim3(1), dim3(64), 0, st, r.ctl.p, r.c->tot.p, mode);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_ctl_init");
    return VNM_OK;
}

// Enqueue one iteration; every kernel reads its sizes and flags from the
// control block, so nothing here waits for the device.
static int lp_run_iter(LPRun &r) {
    vnm_graph *g = r.g;
    LPCache *c = r.c;
    LPRound &w = r.w;
    LPCtl *ctl = r.ctl.p;
    hipStream_t st = g->ctx->stream;
    const int ns = r.ns, M = r.M;
    int rc;
    hipLaunchKernelGGL(k_lp_ctl_begin, dim3(1), dim3(64), 0, st, ctl, r.mode);
    if (r.mode == 0 && (rc = lp_forward(g, c, r.m, 1, &ctl->skip_fwd))) return rc;
    hipLaunchKernelGGL(k_lp_pick, dim3(grid_of(ns, 256)), dim3(256), 0, st, ctl, ns, M, r.mode, r.seed, g->inv.p,
                       c->ptr.p, c->ccls.p, c->segcls.p, c->Y.p, w.dkey.p, w.dkey.p + ns, w.dbest.p, w.abest.p,
                       r.sbits);
    VNM_LAUNCH_CHECK("k_lp_pick");
    if ((rc = sort_keys_u64_if(w.dkey.p, w.sdkey.p, 2 * (int64_t)ns, 33 + r.sbits, st, &ctl->done))) return rc;
    hipLaunchKernelGGL(k_lp_ctl_mid, dim3(1), dim3(64), 0, st, ctl, r.topk, r.window, ns);
    hipLaunchKernelGGL(k_lp_match, dim3(grid_of(r.nnorm, 256)), dim3(256), 0, st, ctl, r.window, ns, M, r.mode,
                       r.seed, w.sdkey.p, w.sdkey.p + ns, w.dbest.p, w.abest.p, g->inv.p, c->Y.p, w.cand.p, r.smask);
    VNM_LAUNCH_CHECK("k_lp_match");
    if ((rc = lp_score_cands(g, c, r.N, w.cand.p, w.best.p, r.ncap, nullptr, &ctl->nc))) return rc;
    hipLaunchKernelGGL(k_lp_gain_keys, dim3(grid_of(r.ncap, 256)), dim3(256), 0, st, ctl, r.ncap, w.cand.p,
                       w.best.p, w.gkey.p, r.ibits, r.gmax);
    // one candidate per donor: the first nnorm keys; the whole window: ncap
    const int kbits = r.gbits + r.ibits;
    if (r.mode != 2 && (rc = sort_keys_u64_if(w.gkey.p, w.sgkey.p, r.nnorm, kbits, st, &ctl->skip_norm))) return rc;
    if (r.mode != 1 && (rc = sort_keys_u64_if(w.gkey.p, w.sgkey.p, r.ncap, kbits, st, &ctl->skip_exh))) return rc;
    const size_t shm = (size_t)((ns + 31) / 32) * 4;
    hipLaunchKernelGGL(k_lp_select, dim3(1), dim3(SEL_T), shm, st, ctl, ns, w.sgkey.p, w.cand.p, w.best.p,
                       w.acc.p, w.owner.p, r.imask);
    VNM_LAUNCH_CHECK("k_lp_select");
    hipLaunchKernelGGL(k_lp_ctl_after, dim3(1), dim3(64), 0, st, ctl, r.mode, g->queue.p);
    const int maxsw = std::min(r.ncap, ns);
    hipLaunchKernelGGL(k_lp_swap, dim3(grid_of(maxsw, 256)), dim3(256), 0, st, ctl, M, w.acc.p, w.cand.p,
                       w.best.p, g->perm.p, g->inv.p, w.touched.p);
    hipLaunchKernelGGL(k_lp_seg_stats, dim3(std::max(1, std::min(2 * maxsw, 1024))), dim3(SS_T), 0, st, g->nrows,
                       M, r.N, ns, w.touched.p, 0, g->inv.p, c->ptr.p, c->idx.p, c->ccls.p, c->segcls.p,
                       c->segnnz.p, g->queue.p, &ctl->nsel, &ctl->skip_swap);
    hipLaunchKernelGGL(k_lp_totals, dim3(1), dim3(1024), 0, st, ns, c->segcls.p, c->tot.p, &ctl->skip_swap);
    hipLaunchKernelGGL(k_lp_ctl_end, dim3(1), dim3(64), 0, st, ctl, c->tot.p);
    g->ctx->launches += 11;
    VNM_LAUNCH_CHECK("k_lp_ctl_end");
    return VNM_OK;
}

static int lp_run_done(LPRun &r, bool *done) {
    int32_t d = 0;
    VNM_HIP_TRY(hipMemcpyAsync(&d, &r.ctl.p->done, 4, hipMemcpyDeviceToHost, r.g->ctx->stream));
    VNM_HIP_TRY(hipStreamSynchronize(r.g->ctx->stream));
    *done = d != 0;
    return VNM_OK;
}

static int lp_run_end(LPRun &r, vnm_lp_stats *stats) {
    hipStream_t st = r.g->ctx->stream;
    int rc = relabel(r.g, r.g->perm.p);
    if (rc) return rc;
    // the copies land in this frame: wait for them on every path
    LPCtl h;
    int32_t tot[2];
    hipError_t e = hipMemcpyAsync(&h, r.ctl.p, sizeof(LPCtl), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(tot, r.c->tot.p, 8, hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return set_error(VNM_E_HIP, "reading the run's results failed: %s", hipGetErrorString(e));
    vnm_lp_stats S{};
    S.init_cnt = r.init_cnt;
    S.init_size = r.init_size;
    S.iters = h.it;
    S.swaps = h.swaps;
    S.gain = h.gain_total;
    S.exact_pairs = h.exact_pairs;
    S.fallbacks = h.fallbacks;
    S.final_cnt = tot[0];
    S.final_size = tot[1];
    S.ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r.t0).count();
    *stats = S;
    return VNM_OK;
}

// The reorder itself, on the stream of g->ctx (arguments checked): one host
// read per iteration, to stop once the run is over.
static int lp_reorder_run(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window,
                          int maxiter, uint64_t seed, vnm_lp_stats *stats) {
    LPRun r;
    int rc = lp_run_begin(r, g, m, N, mode, topk, window, seed);
    if (rc) return rc;
    bool done = false;
    if ((rc = lp_run_done(r, &done))) return rc;
    for (int it = 0; it < maxiter && !done; ++it) {
        if ((rc = lp_run_iter(r))) return rc;
        if ((rc = lp_run_done(r, &done))) return rc;
    }
    return lp_run_end(r, stats);
}

int lp_reorder(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window, int maxiter,
               uint64_t seed, vnm_lp_stats *stats) {
    int rc = lp_reorder_args(g, m, mode, topk, window, maxiter);
    if (rc) return rc;
    return lp_reorder_run(g, m, N, mode, topk, window, maxiter, seed, stats);
}

// The learned reorder of many graphs of one context as one launch chain:
// one iteration of every graph is captured as a HIP graph -- LP_BRANCHES
// parallel branches (forked from the context stream inside the capture),
// the graphs dealt round-robin over them, each branch with its own sort
// scratch; every kernel reads its sizes and flags from its graph's control
// block -- and the graph is replayed until every run is over or maxiter
// iterations have run; the host reads the runs' done flags after each
// replay and captures again with the running graphs once at most half of
// the captured ones still run.  A graph's
// result is the one lp_reorder makes alone with the same seed: the same
// kernels take the same decisions.
constexpr int LP_BRANCHES = 8;

int lp_reorder_batch(vnm_graph *const *gs, int count, const vnm_lp_model *m, int N, int mode, int topk,
                     int window, int maxiter, uint64_t seed, vnm_lp_stats *stats) {
    if (count < 1) return set_error(VNM_E_ARG, "batch of %d graphs", count);
    vnm_ctx *ctx = gs[0] ? gs[0]->ctx : nullptr;
    for (int i = 0; i < count; ++i) {
        if (!gs[i]) return set_error(VNM_E_ARG, "graph %d is null", i);
        if (gs[i]->ctx != ctx) return set_error(VNM_E_ARG, "graphs of a batch share one context");
        int rc = lp_reorder_args(gs[i], m, mode, topk, window, maxiter);
        if (rc) return rc;
        rc = lp_check(gs[i], N);
        if (rc) return rc;
        for (int j = 0; j < i; ++j)
            if (gs[j] == gs[i]) return set_error(VNM_E_ARG, "graph %d appears twice in the batch", i);
    }
    hipStream_t st = ctx->stream;
    const int B = std::min(LP_BRANCHES, count);
    // branch streams (default priority) and fork/join events, released on
    // every path; the graphs get their context back on every path
    struct Res {
        std::vector<vnm_ctx> aux;
        std::vector<hipEvent_t> ev;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        vnm_graph *const *gs = nullptr;
        int count = 0;
        vnm_ctx *ctx = nullptr;
        int slot = 0;
        ~Res() {
            for (int i = 0; i < count; ++i) gs[i]->ctx = ctx;
            (void)sort_scratch_slot(slot);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            for (auto e : ev) (void)hipEventDestroy(e);
            for (auto &x : aux)
                if (x.stream) {
                    (void)hipStreamSynchronize(x.stream);
                    (void)hipStreamDestroy(x.stream);
                }
        }
    } res;
    res.gs = gs;
    res.count = count;
    res.ctx = ctx;
    res.slot = sort_scratch_slot(0);
    res.aux.resize(B);
    for (int b = 0; b < B; ++b) {
        res.aux[b].device = ctx->device;
        hipError_t e = hipStreamCreateWithFlags(&res.aux[b].stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            res.aux[b].stream = nullptr;
            return set_error(VNM_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        }
    }
    for (int k = 0; k < B + 1; ++k) {
        hipEvent_t e = nullptr;
        VNM_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        res.ev.push_back(e);
    }
    // per graph: cache, buffers, control block (on the context stream), sort
    // scratch reserved in its branch's slot
    std::vector<LPRun> runs(count);
    std::vector<LPCtl *> hctl(count);
    int rc;
    for (int i = 0; i < count; ++i) {
        (void)sort_scratch_slot(i % B);
        if ((rc = lp_run_begin(runs[i], gs[i], m, N, mode, topk, window, seed)))
            return set_error(rc, "graph %d: %s", i, last_error());
        hctl[i] = runs[i].ctl.p;
    }
    DBuf<LPCtl *> dctl;
    DBuf<int32_t> dflags;
    if ((rc = dctl.alloc(count)) || (rc = dflags.alloc(count))) return rc;
    VNM_HIP_TRY(hipMemcpyAsync(dctl.p, hctl.data(), count * sizeof(LPCtl *), hipMemcpyHostToDevice, st));
    std::vector<int32_t> done(count, 0);
    VNM_HIP_TRY(hipStreamSynchronize(st));
    // capture one iteration of the graphs in `act` (no event timing inside)
    auto capture = [&](const std::vector<int> &act) -> int {
        if (res.exec) (void)hipGraphExecDestroy(res.exec);
        if (res.graph) (void)hipGraphDestroy(res.graph);
        res.exec = nullptr;
        res.graph = nullptr;
        const bool timing = ctx->timing;
        ctx->timing = false;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            ctx->timing = timing;
            return set_error(VNM_E_HIP, "capture of the learned iteration failed: %s", hipGetErrorString(e));
        }
        int rc = VNM_OK;
        // the branches the running graphs use join the capture (a stream
        // that did not would run its kernels at once instead of recording)
        std::vector<char> used(B, 0);
        for (int i : act) used[i % B] = 1;
        if (hipEventRecord(res.ev[B], st) != hipSuccess) rc = set_error(VNM_E_HIP, "fork event failed");
        for (int b = 0; b < B && rc == VNM_OK; ++b)
            if (used[b] && hipStreamWaitEvent(res.aux[b].stream, res.ev[B], 0) != hipSuccess)
                rc = set_error(VNM_E_HIP, "fork failed");
        for (size_t k = 0; k < act.size() && rc == VNM_OK; ++k) {
            const int i = act[k], b = i % B;  // a graph keeps its branch: its sort scratch slot
            (void)sort_scratch_slot(b);
            gs[i]->ctx = &res.aux[b];
            rc = lp_run_iter(runs[i]);
            gs[i]->ctx = ctx;
            ctx->launches += res.aux[b].launches;
            res.aux[b].launches = 0;
        }
        for (int b = 0; b < B && rc == VNM_OK; ++b)
            if (used[b] && (hipEventRecord(res.ev[b], res.aux[b].stream) != hipSuccess ||
                            hipStreamWaitEvent(st, res.ev[b], 0) != hipSuccess))
                rc = set_error(VNM_E_HIP, "join failed");
        if (rc == VNM_OK) hipLaunchKernelGGL(k_lp_done_flags, dim3(1), dim3(256), 0, st, dctl.p, count, dflags.p);
        e = hipStreamEndCapture(st, &res.graph);
        ctx->timing = timing;
        if (rc) return rc;
        if (e != hipSuccess)
            return set_error(VNM_E_HIP, "capture of the learned iteration failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&res.exec, res.graph, nullptr, nullptr, 0);
        if (e != hipSuccess)
            return set_error(VNM_E_HIP, "instantiating the learned iteration failed: %s", hipGetErrorString(e));
        return VNM_OK;
    };
    // replay; once at most half of the captured graphs still run, capture
    // again with those only (a finished graph's kernels would all return at
    // once, but each still costs a launch)
    std::vector<int> act;
    for (int i = 0; i < count; ++i) act.push_back(i);
    size_t captured = 0;
    const bool dbg = getenv("VNM_DEBUG") != nullptr;
    auto tnow = [] { return std::chrono::steady_clock::now(); };
    auto tms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    auto tb = tnow();
    for (int it = 0; it < maxiter && !act.empty(); ++it) {
        if (2 * act.size() <= captured || captured == 0) {
            auto tc = tnow();
            if ((rc = capture(act))) return rc;
            captured = act.size();
            if (dbg)
                fprintf(stderr, "[vnm] lp batch: iteration %d, capture of %zu graphs %.2f ms (%.2f ms so far)\n", it,
                        captured, tms(tc, tnow()), tms(tb, tnow()));
        }
        VNM_HIP_TRY(hipGraphLaunch(res.exec, st));
        VNM_HIP_TRY(hipMemcpyAsync(done.data(), dflags.p, count * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        std::vector<int> next;
        for (int i : act)
            if (!done[i]) next.push_back(i);
        act.swap(next);
    }
    if (dbg) fprintf(stderr, "[vnm] lp batch: iterations done %.2f ms\n", tms(tb, tnow()));
    for (int i = 0; i < count; ++i)
        if ((rc = lp_run_end(runs[i], &stats[i]))) return set_error(rc, "graph %d: %s", i, last_error());
    return VNM_OK;
}

}  // namespace vnm

using namespace vnm;

// ------------------------------------------------------------ C ABI

extern "C" {

int vnm_lp_model_create(vnm_ctx *ctx, const float *params, uint64_t seed, vnm_lp_model **out) {
    if (!ctx || !out) return set_error(VNM_E_ARG, "bad arguments");
    VNM_HIP_TRY(hipSetDevice(ctx->device));
    vnm_lp_model *m = new vnm_lp_model();
    m->ctx = ctx;
    int rc;
    if ((rc = m->w.alloc(LP_NPARAMS)) || (rc = m->m.alloc(LP_NPARAMS)) || (rc = m->v.alloc(LP_NPARAMS)) ||
        (rc = m->g.alloc(LP_NPARAMS)) || (rc = m->wt.alloc(LP_NWT))) {
        delete m;
        return rc;
    }
    std::vector<float> p(LP_NPARAMS, 0.f);
    if (params) {
        std::copy(params, params + LP_NPARAMS, p.begin());
    } else {
        // He-normal weights (std sqrt(2 / fan_in)), zero biases; the head's
        // 34 used columns N(0, 0.05^2), the rest zero
        std::vector<float> z(LP_NPARAMS);
        vnm_gen_normal(seed, LP_NPARAMS, z.data());
        const int fan[3] = {32, 128, 128};
        for (int l = 0; l < 3; ++l) {
            const float sd = std::sqrt(2.0f / fan[l]);
            for (int i = 0; i < fan[l] * 64; ++i) p[LP_OFF_W[l] + i] = sd * z[LP_OFF_W[l] + i];
        }
        for (int k = 0; k < 64; ++k)
            for (int n = 0; n < 2 + 2 * LP_PQ; ++n) p[LP_OFF_WH + k * 64 + n] = 0.05f * z[LP_OFF_WH + k * 64 + n];
    }
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(m->w.p, p.data(), LP_NPARAMS * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(m->m.p, 0, LP_NPARAMS * 4, st) != hipSuccess ||
        hipMemsetAsync(m->v.p, 0, LP_NPARAMS * 4, st) != hipSuccess ||
        hipMemsetAsync(m->g.p, 0, LP_NPARAMS * 4, st) != hipSuccess) {
        delete m;
        return set_error(VNM_E_HIP, "model upload failed");
    }
    hipLaunchKernelGGL(k_lp_wt, dim3(grid_of(LP_NWT, 256)), dim3(256), 0, st, m->w.p, m->wt.p);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        delete m;
        return set_error(VNM_E_HIP, "model setup failed");
    }
    *out = m;
    return VNM_OK;
}

void vnm_lp_model_destroy(vnm_lp_model *m) {
    if (!m) return;
    if (m->ctx && m->ctx->stream) (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
}

int vnm_lp_model_get(const vnm_lp_model *m, float *params) {
    if (!m || !params) return set_error(VNM_E_ARG, "bad arguments");
    VNM_HIP_TRY(hipMemcpyAsync(params, m->w.p, LP_NPARAMS * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    VNM_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return VNM_OK;
}

int vnm_lp_features(vnm_graph *g, int N, float *X) {
    if (!g || !X) return set_error(VNM_E_ARG, "bad arguments");
    LPCache *c;
    int rc = lp_begin(g, N, &c);
    if (rc) return rc;
    hipStream_t st = g->ctx->stream;
    hipLaunchKernelGGL(k_lp_features<float>, dim3(grid_of(g->nrows, 256)), dim3(256), 0, st, g->nrows, g->M,
                       g->perm.p, c->ptr.p, c->ccls.p, c->segcls.p, c->segnnz.p, c->tot.p, c->X.p, nullptr);
    g->ctx->launches++;
    VNM_LAUNCH_CHECK("k_lp_features");
    VNM_HIP_TRY(hipMemcpyAsync(X, c->X.p, (size_t)g->nrows * LP_F * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int vnm_lp_forward(vnm_lp_model *m, vnm_graph *g, int N, int bf16, float *head) {
    if (!m || !g || !head || m->ctx != g->ctx) return set_error(VNM_E_ARG, "bad arguments");
    LPCache *c;
    int rc = lp_begin(g, N, &c);
    if (rc) return rc;
    if ((rc = lp_forward(g, c, m, bf16 ? 1 : 0))) return rc;
    hipStream_t st = g->ctx->stream;
    VNM_HIP_TRY(hipMemcpyAsync(head, c->Y.p, (size_t)g->nrows * LP_HD * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int vnm_lp_loss_grad(vnm_lp_model *m, vnm_graph *g, int N, int bf16, const int32_t *pairs,
                     const float *target, int count, float *loss, float *grad) {
    if (!m || !g || count < 0 || (count > 0 && (!pairs || !target)) || m->ctx != g->ctx)
        return set_error(VNM_E_ARG, "bad arguments");
    for (int i = 0; i < 2 * count; ++i)
        if (pairs[i] < 0 || pairs[i] >= g->nrows) return set_error(VNM_E_ARG, "pair id out of range");
    LPCache *c;
    int rc = lp_begin(g, N, &c);
    if (rc) return rc;
    if ((rc = lp_forward(g, c, m, bf16 ? 1 : 0))) return rc;
    if ((rc = lp_loss_grad(g, c, m, bf16 ? 1 : 0, pairs, target, count, grad != nullptr, loss))) return rc;
    if (grad) {
        VNM_HIP_TRY(hipMemcpyAsync(grad, m->g.p, LP_NPARAMS * 4, hipMemcpyDeviceToHost, g->ctx->stream));
        VNM_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
    }
    return VNM_OK;
}

int vnm_lp_train_step(vnm_lp_model *m, vnm_graph *g, int N, const int32_t *pairs, const float *target,
                      int count, float lr, float *loss) {
    if (!m || !g || count < 1 || !pairs || !target || !(lr > 0.f) || m->ctx != g->ctx)
        return set_error(VNM_E_ARG, "bad arguments");
    for (int i = 0; i < 2 * count; ++i)
        if (pairs[i] < 0 || pairs[i] >= g->nrows) return set_error(VNM_E_ARG, "pair id out of range");
    LPCache *c;
    int rc = lp_begin(g, N, &c);
    if (rc) return rc;
    if ((rc = lp_forward(g, c, m, 1))) return rc;
    if ((rc = lp_loss_grad(g, c, m, 1, pairs, target, count, true, loss))) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    hipStream_t st = g->ctx->stream;
    hipLaunchKernelGGL(k_lp_adam, dim3(grid_of(LP_NPARAMS, 256)), dim3(256), 0, st, LP_NPARAMS, m->w.p, m->g.p,
                       m->m.p, m->v.p, lr, b1, b2, 1e-8f, c1, c2);
    hipLaunchKernelGGL(k_lp_wt, dim3(grid_of(LP_NWT, 256)), dim3(256), 0, st, m->w.p, m->wt.p);
    g->ctx->launches += 2;
    VNM_LAUNCH_CHECK("k_lp_adam");
    return VNM_OK;
}

int vnm_pair_score(vnm_graph *g, int N, const int32_t *segp, const int32_t *segq, int count, int32_t *scores) {
    if (!g || count < 0 || (count > 0 && (!segp || !segq || !scores))) return set_error(VNM_E_ARG, "bad arguments");
    LPCache *c;
    int rc = lp_check(g, N);
    if (rc) return rc;
    const int ns = g->nseg();
    std::vector<Cand> hc(count);
    for (int i = 0; i < count; ++i) {
        if (segp[i] < 0 || segp[i] >= ns || segq[i] < 0 || segq[i] >= ns)
            return set_error(VNM_E_ARG, "segment out of range at %d", i);
        if (segp[i] == segq[i]) return set_error(VNM_E_ARG, "pair %d scores a segment against itself", i);
        hc[i] = Cand{segp[i], segq[i], 0, 0};
    }
    if (count == 0) return VNM_OK;
    if ((rc = lp_begin(g, N, &c))) return rc;
    DBuf<Cand> dc;
    DBuf<int4> db;
    DBuf<int32_t> df;
    const int MM = g->M * g->M;
    if ((rc = dc.alloc(count)) || (rc = db.alloc(count)) || (rc = df.alloc((size_t)count * 2 * MM))) return rc;
    hipStream_t st = g->ctx->stream;
    VNM_HIP_TRY(hipMemcpyAsync(dc.p, hc.data(), (size_t)count * sizeof(Cand), hipMemcpyHostToDevice, st));
    if ((rc = lp_score_cands(g, c, N, dc.p, db.p, count, df.p))) return rc;
    VNM_HIP_TRY(hipMemcpyAsync(scores, df.p, (size_t)count * 2 * MM * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int vnm_lp_reorder(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window, int maxiter,
                   uint64_t seed, vnm_lp_stats *st) {
    if (!g || !st) return set_error(VNM_E_ARG, "bad arguments");
    return lp_reorder(g, m, N, mode, topk, window, maxiter, seed, st);
}

int vnm_lp_reorder_batch(vnm_graph *const *graphs, int count, const vnm_lp_model *m, int N, int mode, int topk,
                         int window, int maxiter, uint64_t seed, vnm_lp_stats *st) {
    if (!graphs || !st) return set_error(VNM_E_ARG, "bad arguments");
    return lp_reorder_batch(graphs, count, m, N, mode, topk, window, maxiter, seed, st);
}

}  // extern "C"
