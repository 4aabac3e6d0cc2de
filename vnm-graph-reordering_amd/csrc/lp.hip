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
    int32_t *__restrict__ queue) {
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
                                                     int32_t *__restrict__ tot) {
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
                              const int32_t *__restrict__ tot, T *__restrict__ X) {
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
                                                      const TI *__restrict__ H, TO *__restrict__ out) {
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
                                                       const TI *__restrict__ H, float *__restrict__ part) {
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nch) return;
    const float4 acc = agg_range<TI, F, BWD>(ch_beg[c], ch_end[c], idx, inv_deg, H);
    const int lane = threadIdx.x & 63;
    if (lane < F / 4) *(float4 *)&part[(size_t)c * F + lane * 4] = acc;
}

template <typename TO, int F, bool BWD>
__global__ void k_lp_agg_combine(int nheavy, const int32_t *__restrict__ heavy,
                                 const int32_t *__restrict__ ch0, const float *__restrict__ part,
                                 const float *__restrict__ inv_deg, TO *__restrict__ out) {
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
                                                       TO *__restrict__ out) {
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
                                                      float *__restrict__ out) {
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

// Per segment: the donor (a node of an invalid segment that sits in a row
// with N+1 ones, best donor score, ties to the lower bit) and the acceptor
// (best acceptor score).  mode 0: network scores Y; 1: seeded hash scores;
// 2: counts (donor = its rows with N+1 ones; acceptor = minus its degree and
// its segment's rows with at least N ones).  Keys sort ascending = best first.
__global__ void k_lp_pick(int nseg, int M, int mode, uint64_t seed, const int32_t *__restrict__ inv,
                          const int32_t *__restrict__ ptr, const int32_t *__restrict__ ccls,
                          const int32_t *__restrict__ segcls, const float *__restrict__ Y,
                          uint64_t *__restrict__ dkey, uint64_t *__restrict__ akey,
                          int8_t *__restrict__ dbest, int8_t *__restrict__ abest,
                          int32_t *__restrict__ ndonor, int sbits) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
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
    dkey[s] = bdi >= 0 ? ((uint64_t)(~ord_f32(bd)) << sbits) | (uint32_t)s : ~0ull;
    akey[s] = ((uint64_t)(~ord_f32(ba)) << sbits) | (uint32_t)s;
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
__global__ void k_lp_match(int nd, int W, int nseg, int stride, int M, int mode, int all, uint64_t seed,
                           const uint64_t *__restrict__ sdkey, const uint64_t *__restrict__ sakey,
                           const int8_t *__restrict__ dbest, const int8_t *__restrict__ abest,
                           const int32_t *__restrict__ inv, const float *__restrict__ Y,
                           Cand *__restrict__ cand, uint64_t smask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
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
__global__ __launch_bounds__(256) void k_lp_pair_gain(int R, int M, int N, const Cand *__restrict__ cand,
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
__global__ void k_lp_gain_keys(int nc, const Cand *__restrict__ cand, const int4 *__restrict__ best,
                               uint64_t *__restrict__ keys, int ibits, uint32_t gmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int4 b = best[i];
    const int gain = b.z + b.w;
    keys[i] = (cand[i].q >= 0 && gain > 0) ? ((uint64_t)(gmax - (uint32_t)gain) << ibits) | (uint32_t)i : ~0ull;
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
__global__ __launch_bounds__(SEL_T) void k_lp_select(int nc, int nseg, const uint64_t *__restrict__ skeys,
                                                     const Cand *__restrict__ cand,
                                                     const int4 *__restrict__ best, int32_t *__restrict__ acc,
                                                     int32_t *__restrict__ out, int32_t *owner, uint64_t imask) {
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
    if (tid == 0) { out[0] = nacc; out[1] = gsum; }
}

// Apply the accepted swaps (segment-disjoint, so in parallel) and list the
// touched segments.
__global__ void k_lp_swap(int nsw, int M, const int32_t *__restrict__ acc, const Cand *__restrict__ cand,
                          const int4 *__restrict__ best, int32_t *__restrict__ perm,
                          int32_t *__restrict__ inv, int32_t *__restrict__ touched) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsw) return;
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
                       g->inv.p, c->ptr.p, c->idx.p, c->ccls.p, c->segcls.p, c->segnnz.p, g->queue.p);
    hipLaunchKernelGGL(k_lp_totals, dim3(1), dim3(1024), 0, st, c->nseg, c->segcls.p, c->tot.p);
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
static int lp_aggregate(vnm_graph *g, LPCache *c, const TI *H, TO *out) {
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
                       c->inv_deg.p, H, out);
    g->ctx->launches++;
    if (P.nch) {
        hipLaunchKernelGGL((k_lp_agg_chunks<TI, F, BWD>), dim3((P.nch + 3) / 4), dim3(256), 0, st, P.nch,
                           P.beg.p, P.end.p, idx, c->inv_deg.p, H, P.part.p);
        hipLaunchKernelGGL((k_lp_agg_combine<TO, F, BWD>), dim3(grid_of((int64_t)P.nheavy * F, 256)),
                           dim3(256), 0, st, P.nheavy, P.heavy.p, P.ch0.p, P.part.p, c->inv_deg.p, out);
        g->ctx->launches += 2;
    }
    VNM_LAUNCH_CHECK("k_lp_agg");
    return VNM_OK;
}

// Features, three layers and the head at the current perm/inv; Y is FP32.
static int lp_forward(vnm_graph *g, LPCache *c, const vnm_lp_model *m, int bf16) {
    hipStream_t st = g->ctx->stream;
    const int R = c->R;
    int rc;
    KernelTimer kt(g->ctx, "lp_forward");
    if (bf16) {
        typedef uint16_t T;
        T *X = (T *)c->X.p, *A0 = (T *)c->A0.p, *H1 = (T *)c->H1.p, *A1 = (T *)c->A1.p;
        T *H2 = (T *)c->H2.p, *A2 = (T *)c->A2.p, *H3 = (T *)c->H3.p;
        hipLaunchKernelGGL(k_lp_features<T>, dim3(grid_of(R, 256)), dim3(256), 0, st, R, g->M, g->perm.p,
                           c->ptr.p, c->ccls.p, c->segcls.p, c->segnnz.p, c->tot.p, X);
        if ((rc = lp_aggregate<T, T, 16, false>(g, c, X, A0))) return rc;
        const int gr = (R + 63) / 64;
        hipLaunchKernelGGL((k_lp_dense_mfma<16, 16, true, T>), dim3(gr), dim3(256), 0, st, R, X, A0,
                           m->wt.p + LP_WT_OFF[0], m->w.p + LP_OFF_B[0], H1);
        if ((rc = lp_aggregate<T, T, 64, false>(g, c, H1, A1))) return rc;
        hipLaunchKernelGGL((k_lp_dense_mfma<64, 64, true, T>), dim3(gr), dim3(256), 0, st, R, H1, A1,
                           m->wt.p + LP_WT_OFF[1], m->w.p + LP_OFF_B[1], H2);
        if ((rc = lp_aggregate<T, T, 64, false>(g, c, H2, A2))) return rc;
        hipLaunchKernelGGL((k_lp_dense_mfma<64, 64, true, T>), dim3(gr), dim3(256), 0, st, R, H2, A2,
                           m->wt.p + LP_WT_OFF[2], m->w.p + LP_OFF_B[2], H3);
        hipLaunchKernelGGL((k_lp_dense_mfma<64, 0, false, float>), dim3(gr), dim3(256), 0, st, R, H3,
                           (const T *)nullptr, m->wt.p + LP_WT_OFF[3], m->w.p + LP_OFF_BH, c->Y.p);
    } else {
        float *X = c->X.p, *A0 = c->A0.p, *H1 = c->H1.p, *A1 = c->A1.p, *H2 = c->H2.p, *A2 = c->A2.p,
              *H3 = c->H3.p;
        hipLaunchKernelGGL(k_lp_features<float>, dim3(grid_of(R, 256)), dim3(256), 0, st, R, g->M,
                           g->perm.p, c->ptr.p, c->ccls.p, c->segcls.p, c->segnnz.p, c->tot.p, X);
        const int gr = grid_of(R, 4, 4096);
        if ((rc = lp_aggregate<float, float, 16, false>(g, c, X, A0))) return rc;
        hipLaunchKernelGGL((k_lp_dense_f32<16, 16, true>), dim3(gr), dim3(256), 0, st, R, X, A0,
                           m->w.p + LP_OFF_W[0], m->w.p + LP_OFF_B[0], H1);
        if ((rc = lp_aggregate<float, float, 64, false>(g, c, H1, A1))) return rc;
        hipLaunchKernelGGL((k_lp_dense_f32<64, 64, true>), dim3(gr), dim3(256), 0, st, R, H1, A1,
                           m->w.p + LP_OFF_W[1], m->w.p + LP_OFF_B[1], H2);
        if ((rc = lp_aggregate<float, float, 64, false>(g, c, H2, A2))) return rc;
        hipLaunchKernelGGL((k_lp_dense_f32<64, 64, true>), dim3(gr), dim3(256), 0, st, R, H2, A2,
                           m->w.p + LP_OFF_W[2], m->w.p + LP_OFF_B[2], H3);
        hipLaunchKernelGGL((k_lp_dense_f32<64, 0, false>), dim3(gr), dim3(256), 0, st, R, H3,
                           (const float *)nullptr, m->w.p + LP_OFF_WH, m->w.p + LP_OFF_BH, c->Y.p);
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
    DBuf<uint64_t> dkey, akey, sdkey, sakey, gkey, sgkey;
    DBuf<int8_t> dbest, abest;
    DBuf<Cand> cand;
    DBuf<int4> best;
    DBuf<int32_t> acc, out, touched, owner;
};

// Score candidates [0, nc) exactly: a workgroup per candidate.
static int lp_score_cands(vnm_graph *g, LPCache *c, int N, Cand *cand, int4 *best, int nc, int32_t *full) {
    if (nc <= 0) return VNM_OK;
    KernelTimer kt(g->ctx, "k_lp_pair_gain");
    hipLaunchKernelGGL(k_lp_pair_gain, dim3(nc), dim3(256), 0, g->ctx->stream, g->nrows, g->M, N, cand,
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

// The reorder itself, on the stream of g->ctx (arguments checked).
static int lp_reorder_run(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window,
                          int maxiter, uint64_t seed, vnm_lp_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    LPCache *c;
    int rc = lp_begin(g, N, &c);
    if (rc) return rc;
    hipStream_t st = g->ctx->stream;
    const int ns = c->nseg, M = g->M;
    const int ncmax = std::max(topk * window, topk);
    // sort keys hold a segment or candidate index in their low bits
    auto bits_of = [](int64_t x) { int b = 1; while ((1ll << b) <= x) ++b; return b; };
    const int sbits = bits_of(ns - 1), ibits = bits_of(ncmax - 1), gbits = bits_of(2 * (int64_t)g->nrows);
    const uint32_t gmax = (uint32_t)((1ull << gbits) - 1);
    const uint64_t smask = (1ull << sbits) - 1, imask = (1ull << ibits) - 1;
    LPRound w;
    if ((rc = w.dkey.alloc(ns)) || (rc = w.akey.alloc(ns)) || (rc = w.sdkey.alloc(ns)) ||
        (rc = w.sakey.alloc(ns)) || (rc = w.dbest.alloc(ns)) || (rc = w.abest.alloc(ns)) ||
        (rc = w.cand.alloc(ncmax)) || (rc = w.best.alloc(ncmax)) || (rc = w.gkey.alloc(ncmax)) ||
        (rc = w.sgkey.alloc(ncmax)) || (rc = w.acc.alloc(ncmax)) || (rc = w.out.alloc(8)) ||
        (rc = w.touched.alloc(2 * (size_t)ncmax)) || (rc = w.owner.alloc(ns)))
        return rc;
    VNM_HIP_TRY(hipMemsetAsync(w.owner.p, 0x7F, (size_t)ns * 4, st));  // OWNER_NONE in every slot
    int32_t tot[2];
    VNM_HIP_TRY(hipMemcpyAsync(tot, c->tot.p, 8, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    vnm_lp_stats S{};
    S.init_cnt = tot[0];
    S.init_size = tot[1];
    bool exhaustive_now = (mode == 2);
    int it = 0;
    while (it < maxiter && tot[0] > 0) {
        const int pick_mode = exhaustive_now ? 2 : mode;
        if (pick_mode == 0 && (rc = lp_forward(g, c, m, 1))) return rc;
        VNM_HIP_TRY(hipMemsetAsync(w.out.p, 0, 32, st));
        hipLaunchKernelGGL(k_lp_pick, dim3(grid_of(ns, 256)), dim3(256), 0, st, ns, M, pick_mode,
                           seed + (uint64_t)it * 0x9E3779B97F4A7C15ull, g->inv.p, c->ptr.p, c->ccls.p,
                           c->segcls.p, c->Y.p, w.dkey.p, w.akey.p, w.dbest.p, w.abest.p, w.out.p + 4, sbits);
        g->ctx->launches++;
        VNM_LAUNCH_CHECK("k_lp_pick");
        if ((rc = sort_keys_u64(w.dkey.p, w.sdkey.p, ns, 32 + sbits, st))) return rc;
        if ((rc = sort_keys_u64(w.akey.p, w.sakey.p, ns, 32 + sbits, st))) return rc;
        int32_t nd_h = 0;
        VNM_HIP_TRY(hipMemcpyAsync(&nd_h, w.out.p + 4, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        const int nd = std::min(topk, nd_h);
        if (nd == 0) break;
        const int all = exhaustive_now ? 1 : 0;
        const int W = window;
        const int stride = std::max(1, ns / std::max(nd, 1));
        hipLaunchKernelGGL(k_lp_match, dim3(grid_of(nd, 256)), dim3(256), 0, st, nd, W, ns, stride, M, pick_mode,
                           all, seed ^ ((uint64_t)it << 20), w.sdkey.p, w.sakey.p, w.dbest.p, w.abest.p, g->inv.p,
                           c->Y.p, w.cand.p, smask);
        g->ctx->launches++;
        VNM_LAUNCH_CHECK("k_lp_match");
        const int nc = all ? nd * W : nd;
        if ((rc = lp_score_cands(g, c, N, w.cand.p, w.best.p, nc, nullptr))) return rc;
        S.exact_pairs += nc;
        hipLaunchKernelGGL(k_lp_gain_keys, dim3(grid_of(nc, 256)), dim3(256), 0, st, nc, w.cand.p, w.best.p,
                           w.gkey.p, ibits, gmax);
        if ((rc = sort_keys_u64(w.gkey.p, w.sgkey.p, nc, gbits + ibits, st))) return rc;
        const size_t shm = (size_t)((ns + 31) / 32) * 4;
        hipLaunchKernelGGL(k_lp_select, dim3(1), dim3(SEL_T), shm, st, nc, ns, w.sgkey.p, w.cand.p, w.best.p,
                           w.acc.p, w.out.p, w.owner.p, imask);
        g->ctx->launches += 2;
        VNM_LAUNCH_CHECK("k_lp_select");
        int32_t o[2];
        VNM_HIP_TRY(hipMemcpyAsync(o, w.out.p, 8, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        ++it;
        if (o[0] == 0) {
            if (mode == 0 && !exhaustive_now) {
                // nothing the proposals found pays: score the whole window exactly
                exhaustive_now = true;
                S.fallbacks++;
                continue;
            }
            break;
        }
        hipLaunchKernelGGL(k_lp_swap, dim3(grid_of(o[0], 256)), dim3(256), 0, st, o[0], M, w.acc.p, w.cand.p,
                           w.best.p, g->perm.p, g->inv.p, w.touched.p);
        g->ctx->launches++;
        VNM_LAUNCH_CHECK("k_lp_swap");
        if ((rc = lp_stats(g, c, N, w.touched.p, 2 * o[0]))) return rc;
        VNM_HIP_TRY(hipMemcpyAsync(tot, c->tot.p, 8, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        S.swaps += o[0];
        S.gain += o[1];
        if (mode == 0) exhaustive_now = false;
    }
    S.iters = it;
    S.final_cnt = tot[0];
    S.final_size = tot[1];
    if ((rc = relabel(g, g->perm.p))) return rc;
    VNM_HIP_TRY(hipStreamSynchronize(st));
    S.ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *stats = S;
    return VNM_OK;
}

int lp_reorder(vnm_graph *g, const vnm_lp_model *m, int N, int mode, int topk, int window, int maxiter,
               uint64_t seed, vnm_lp_stats *stats) {
    int rc = lp_reorder_args(g, m, mode, topk, window, maxiter);
    if (rc) return rc;
    return lp_reorder_run(g, m, N, mode, topk, window, maxiter, seed, stats);
}

// The learned reorder of many graphs of one context: `workers` host threads,
// each with a stream of its own, take the graphs in turn (a graph's run
// launches dozens of small kernels and waits on the host three times per
// iteration, so runs on separate streams overlap).  A graph's run, and so
// its result, is the one lp_reorder makes alone with the same seed.
int lp_reorder_batch(vnm_graph *const *gs, int count, const vnm_lp_model *m, int N, int mode, int topk,
                     int window, int maxiter, uint64_t seed, int workers, vnm_lp_stats *stats) {
    if (count < 1) return set_error(VNM_E_ARG, "batch of %d graphs", count);
    if (workers < 1 || workers > 16) return set_error(VNM_E_ARG, "workers must be 1..16 (got %d)", workers);
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
    VNM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int W = std::min(workers, count);
    // one stream per worker, default priority, released on every path
    struct Aux {
        std::vector<vnm_ctx> c;
        explicit Aux(int n) : c(n) {}
        ~Aux() {
            for (auto &x : c)
                if (x.stream) {
                    (void)hipStreamSynchronize(x.stream);
                    (void)hipStreamDestroy(x.stream);
                }
        }
    } aux(W);
    for (int w = 0; w < W; ++w) {
        aux.c[w].device = ctx->device;
        hipError_t e = hipStreamCreateWithFlags(&aux.c[w].stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            aux.c[w].stream = nullptr;
            return set_error(VNM_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        }
    }
    std::atomic<int> next{0}, failed{0};
    std::vector<int> wrc(W, VNM_OK), wat(W, -1);
    std::vector<std::string> wmsg(W);
    auto work = [&](int w) {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            wrc[w] = VNM_E_HIP;
            wmsg[w] = "hipSetDevice failed in a worker";
            failed.store(1);
            return;
        }
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= count || failed.load()) return;
            vnm_graph *g = gs[i];
            g->ctx = &aux.c[w];
            const int rc = lp_reorder_run(g, m, N, mode, topk, window, maxiter, seed, &stats[i]);
            g->ctx = ctx;
            if (rc) {
                wrc[w] = rc;
                wat[w] = i;
                wmsg[w] = last_error();
                failed.store(1);
                return;
            }
        }
    };
    std::vector<std::thread> th;
    try {  // workers that cannot be started leave their graphs to the others
        for (int w = 1; w < W; ++w) th.emplace_back(work, w);
    } catch (const std::system_error &) {
    }
    work(0);
    for (auto &t : th) t.join();
    for (auto &x : aux.c) ctx->launches += x.launches;
    // the failure of the lowest graph index is reported
    int bad = -1;
    for (int w = 0; w < W; ++w)
        if (wrc[w] != VNM_OK && (bad < 0 || wat[w] < wat[bad] || wat[bad] < 0)) bad = w;
    if (bad >= 0) return set_error(wrc[bad], "graph %d: %s", wat[bad], wmsg[bad].c_str());
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
                       g->perm.p, c->ptr.p, c->ccls.p, c->segcls.p, c->segnnz.p, c->tot.p, c->X.p);
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
                         int window, int maxiter, uint64_t seed, int workers, vnm_lp_stats *st) {
    if (!graphs || !st) return set_error(VNM_E_ARG, "bad arguments");
    return lp_reorder_batch(graphs, count, m, N, mode, topk, window, maxiter, seed, workers, st);
}

}  // extern "C"
