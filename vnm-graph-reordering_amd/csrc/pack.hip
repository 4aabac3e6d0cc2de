// V:N:M (N = 2) packer and SpMM for gfx950.
//
// Replaces the reference's SpMM paths (kernel/spmm.cu: the 16x16 WMMA BSR
// kernel :44-161 and the cuSPARSELt 2:4 prune/compress/matmul :324-500,
// which densify A first, kernel/eval.h:12-18).
//
// Format (VENOM with row compression).  A is cut into V x M blocks (V rows,
// one segment of M columns).  Each stored block keeps 4 of its M columns --
// the subset maximising sum over rows of min(2, |row & S|), ties to the
// lexicographically smallest subset -- and each row keeps at most 2 values
// inside them (2:4): the two largest |v|, ties to the lower column.  Blocks
// store their segment, the 4 column offsets (4 bits each) and, for FP8, one
// FP32 scale (absmax / 448).  Rows of a block that keep anything store one
// "row entry": two values and two 2-bit selector indices.  Row entries are
// laid out row-major (CSR over rows), so the SpMM accumulates each output row
// in registers and writes it once.
//
//   bytes(A) = 4 (R + 1) + E (4 + 4 + 1 + 2 s_v) + nb (4 + 2 + 4)
// (rowptr; per row entry its row, block, selector byte and two values; per
// block its segment, selectors and scale).
//
// SpMM: C[R x n] = A_vnm B, FP32 accumulation.  Packing also writes the
// SpMM's value stream: every kept value of the row entries, row-major, as
// (row, B row, FP32 weight = value x block scale), 12 bytes per value, so the
// SpMM reads no block data and branches on no selector.  A work item holds
// whole rows of at most 64 values, or a 64-value chunk of a longer row (item
// table built once per packed matrix); a wave takes one item for one column
// tile of 64 * VEC columns and walks it row by row: the row's B rows and
// weights arrive as scalar loads, every lane gathers VEC consecutive values of
// each B row (2 to 16 bytes) into VEC FMAs with the weight as a scalar
// operand, and owns VEC columns of the output row, which is stored when the
// row ends (rows that keep nothing are written in the gaps).  The chunks of a
// long row store partial sums into slots, and a fix-up pass adds each cut
// row's slots in item order, so C is bit-identical for every launch shape and
// run: no atomics.  The T column tiles interleave in 128-byte pieces of each
// row, and when T divides 8, tile t runs on its own 8/T XCDs, so each XCD's L2
// serves one 1/T slice of B spread over all of its channels.  VEC is chosen per (B
// dtype, n) by timing the usable values on first use (vnm_spmm_set_shape
// overrides).  FP8 B is decoded by the gfx950 e4m3 conversion instruction.
#include "graph.h"
#include "hip_common.h"
#include "packed.h"
#include "sortlib.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace vnm {

static int grid_of(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
}

static int bits_for(int64_t x) {
    int b = 1;
    while ((1ll << b) <= x) ++b;
    return b;
}

// ------------------------------------------------------------ value codecs

// ------------------------------------------------------------ packer kernels

__global__ void k_nz_keys(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                          const float *__restrict__ val, int64_t nnz, int32_t ncolseg,
                          uint64_t *__restrict__ keys, int32_t *__restrict__ idx) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const bool ok = val[i] != 0.0f && col[i] < ncolseg;
        keys[i] = ok ? (((uint64_t)(uint32_t)row[i] << 32) | (uint32_t)col[i]) : ~0ull;
        idx[i] = (int32_t)i;
    }
}

__global__ void k_pk_heads(const uint64_t *__restrict__ k, int64_t n, int M, int32_t *__restrict__ head) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int h = 1;
        if (i > 0) {
            uint64_t a = k[i], b = k[i - 1];
            h = ((a >> 32) != (b >> 32)) || ((uint32_t)a / M != (uint32_t)b / M);
        }
        head[i] = h;
    }
}

__global__ void k_pk_runs(const uint64_t *__restrict__ k, const int32_t *__restrict__ head,
                          const int32_t *__restrict__ excl, int64_t n, int M, int V, int nblk,
                          uint32_t *__restrict__ bits, int32_t *__restrict__ rstart,
                          int32_t *__restrict__ rlen, uint64_t *__restrict__ bkey,
                          int32_t *__restrict__ rid) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = k[i];
        const uint32_t c = (uint32_t)key, r = (uint32_t)(key >> 32);
        const int u = excl[i] + head[i] - 1;
        atomicOr(&bits[u], 1u << (M - 1 - (c % M)));
        atomicAdd(&rlen[u], 1);
        if (head[i]) {
            rstart[u] = (int32_t)i;
            // block key: (segment, row / V), then the row (stable order)
            bkey[u] = ((uint64_t)((uint64_t)(c / M) * nblk + r / V) << 32) | r;
            rid[u] = u;
        }
    }
}

__global__ void k_blk_heads(const uint64_t *__restrict__ bk, int U, int32_t *__restrict__ head) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < U) head[i] = (i == 0) || ((bk[i] >> 32) != (bk[i - 1] >> 32));
}

// The selected subset's 4 column offsets (ascending), 4 bits each.
__device__ __forceinline__ uint16_t subset_sel(uint32_t S, int M) {
    uint16_t sel = 0;
    int k = 0;
    for (int c = 0; c < M && k < 4; ++c)
        if (S & (1u << (M - 1 - c))) sel |= (uint16_t)(c << (4 * k++));
    return sel;
}

constexpr int LANE_RUNS = 8;  // runs of a block that one lane scores

// Block selection, one lane per block of at most LANE_RUNS runs (nearly all
// blocks of sparse graphs): the run bitmaps stay in registers and the
// subsets are scored in ascending index order with strict > (ties to the
// smallest), the order k_blk_select's argmax yields.  Longer blocks are
// listed for k_blk_select.
__global__ __launch_bounds__(256) void k_blk_select_lane(
    const uint64_t *__restrict__ bk, const int32_t *__restrict__ rid_sorted, int U, int nb, int M, int nblk,
    const uint32_t *__restrict__ bits, const uint32_t *__restrict__ subs, int nsub,
    const int32_t *__restrict__ blk_first, int32_t *__restrict__ blk_seg, uint16_t *__restrict__ blk_sel,
    uint32_t *__restrict__ blk_mask, int32_t *__restrict__ run_blk, unsigned long long *__restrict__ retained,
    int32_t *__restrict__ longb, int32_t *__restrict__ nlong) {
    int kept = 0;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const int first = blk_first[b];
        const int last = b + 1 < nb ? blk_first[b + 1] : U;
        const int n = last - first;
        if (n > LANE_RUNS) {
            longb[atomicAdd(nlong, 1)] = b;
        } else {
            uint32_t r[LANE_RUNS];
#pragma unroll
            for (int j = 0; j < LANE_RUNS; ++j) r[j] = j < n ? bits[rid_sorted[first + j]] : 0u;
            int best_score = -1, best_idx = 0;
            for (int s = 0; s < nsub; ++s) {
                const uint32_t S = subs[s];
                int sc = 0;
#pragma unroll
                for (int j = 0; j < LANE_RUNS; ++j) sc += min(2, popc32(r[j] & S));
                if (sc > best_score) {
                    best_score = sc;
                    best_idx = s;
                }
            }
            const uint32_t S = subs[best_idx];
            blk_seg[b] = (int32_t)((bk[first] >> 32) / (uint64_t)nblk);
            blk_sel[b] = subset_sel(S, M);
            blk_mask[b] = S;
            kept += best_score;
            for (int j = 0; j < n; ++j) run_blk[rid_sorted[first + j]] = b;
        }
    }
    // one atomic per workgroup: atomics on a single address serialise at the
    // L2, so the grid is capped and the workgroup sums first
    __shared__ int wsum[4];
    for (int d = 32; d >= 1; d >>= 1) kept += __shfl_xor(kept, d, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(retained, (unsigned long long)t);
    }
}

// One wave per listed block: best 4-column subset (lexicographic order of
// ascending column tuples; ties to the smallest).  nsub = C(M, 4) <= 70.
__global__ __launch_bounds__(256) void k_blk_select(const uint64_t *__restrict__ bk, const int32_t *__restrict__ rid_sorted,
                             int U, int nb, int M, int nblk, const uint32_t *__restrict__ bits,
                             const uint32_t *__restrict__ subs, int nsub,
                             const int32_t *__restrict__ blk_first, int32_t *__restrict__ blk_seg,
                             uint16_t *__restrict__ blk_sel, uint32_t *__restrict__ blk_mask,
                             int32_t *__restrict__ run_blk, unsigned long long *__restrict__ retained,
                             const int32_t *__restrict__ list, const int32_t *__restrict__ nlist) {
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x / 64;
    const int count = *nlist;
    // listed blocks in a grid-stride loop per wave; the retained count is
    // summed per wave and added once (one atomic per block on a single
    // address serialises at the L2)
    unsigned long long kept = 0;
    for (int i = blockIdx.x * wpb + (threadIdx.x >> 6); i < count; i += gridDim.x * wpb) {
        const int b = list[i];
        const int first = blk_first[b];
        int last = U;
        if (b + 1 < nb) last = blk_first[b + 1];
        // scores of 128 subsets at a time (lane and lane + 64; one round for
        // M <= 8): the block's run bitmaps are loaded 64 at a time, one per
        // lane, and broadcast
        int best_score = -1, best_idx = 0x7FFFFFFF;
        for (int s0 = 0; s0 < nsub; s0 += 128) {
            const int i0 = s0 + lane, i1 = s0 + 64 + lane;
            const uint32_t S0 = i0 < nsub ? subs[i0] : 0u;
            const uint32_t S1 = i1 < nsub ? subs[i1] : 0u;
            int sc0 = 0, sc1 = 0;
            for (int base = first; base < last; base += 64) {
                const int q = base + lane;
                const int mine = q < last ? (int)bits[rid_sorted[q]] : 0;
                const int n = min(64, last - base);
                for (int j = 0; j < n; ++j) {
                    const uint32_t bj = (uint32_t)__builtin_amdgcn_readlane(mine, j);
                    sc0 += min(2, popc32(bj & S0));
                    sc1 += min(2, popc32(bj & S1));
                }
            }
            // ascending subset index per lane: strict > keeps the smaller on ties
            if (i0 < nsub && sc0 > best_score) { best_score = sc0; best_idx = i0; }
            if (i1 < nsub && sc1 > best_score) { best_score = sc1; best_idx = i1; }
        }
        // wave argmax (score desc, index asc)
        for (int d = 32; d >= 1; d >>= 1) {
            int os = __shfl_xor(best_score, d, 64), oi = __shfl_xor(best_idx, d, 64);
            if (os > best_score || (os == best_score && oi < best_idx)) { best_score = os; best_idx = oi; }
        }
        const uint32_t S = subs[best_idx];
        if (lane == 0) {
            const uint64_t key = bk[first] >> 32;
            blk_seg[b] = (int32_t)(key / (uint64_t)nblk);
            blk_sel[b] = subset_sel(S, M);
            blk_mask[b] = S;
            kept += (unsigned long long)best_score;
        }
        for (int q = first + lane; q < last; q += 64) run_blk[rid_sorted[q]] = b;
    }
    // one atomic per workgroup (lane 0 of each wave holds the wave's sum)
    __shared__ unsigned long long wsum[4];
    if (lane == 0) wsum[threadIdx.x >> 6] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(retained, t);
    }
}

__global__ void k_blk_first(const int32_t *__restrict__ head, const int32_t *__restrict__ excl, int U,
                            int32_t *__restrict__ blk_first) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < U && head[i]) blk_first[excl[i]] = i;
}

// Per run: the kept values (<= 2) among the block's selected columns.
__global__ void k_run_keep(const uint64_t *__restrict__ keys, const int32_t *__restrict__ eidx,
                           const float *__restrict__ val, int U, int M,
                           const int32_t *__restrict__ rstart, const int32_t *__restrict__ rlen,
                           const int32_t *__restrict__ run_blk, const uint32_t *__restrict__ blk_mask,
                           float *__restrict__ kv, uint8_t *__restrict__ kidx,
                           int32_t *__restrict__ nkeep, uint32_t *__restrict__ blk_absmax) {
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= U) return;
    const int b = run_blk[u];
    const uint32_t S = blk_mask[b];
    // selector index of column offset c within S (ascending order)
    float v0 = 0.f, v1 = 0.f;
    int c0 = -1, c1 = -1;
    const int beg = rstart[u], len = rlen[u];
    for (int q = beg; q < beg + len; ++q) {
        const int c = (int)((uint32_t)keys[q] % (uint32_t)M);
        if (!(S & (1u << (M - 1 - c)))) continue;
        const float v = val[eidx[q]];
        const float a = fabsf(v);
        // entries come in ascending column order: strict > keeps the lower
        // column on ties
        if (c0 < 0 || a > fabsf(v0)) { v1 = v0; c1 = c0; v0 = v; c0 = c; }
        else if (c1 < 0 || a > fabsf(v1)) { v1 = v; c1 = c; }
    }
    int n = (c0 >= 0) + (c1 >= 0);
    uint8_t code = 0;
    if (n > 0) {
        int i0 = popc32(S >> (M - c0)), i1 = c1 >= 0 ? popc32(S >> (M - c1)) : 0;
        code = (uint8_t)(i0 | (i1 << 2) | (c1 >= 0 ? 0x10 : 0));
        float m = fmaxf(fabsf(v0), c1 >= 0 ? fabsf(v1) : 0.f);
        atomicMax(&blk_absmax[b], __float_as_uint(m));
    }
    kv[2 * u] = v0;
    kv[2 * u + 1] = c1 >= 0 ? v1 : 0.f;
    kidx[u] = code;
    nkeep[u] = n;
}

// Kept runs per row.  Runs are sorted by row, so a wave's lanes of one row
// are contiguous: the last lane of each such segment adds the segment's
// kept count (one atomic per row and wave; a hub row's thousands of runs
// would otherwise serialise on one address).
__global__ void k_row_counts(const uint64_t *__restrict__ keys, const int32_t *__restrict__ rstart,
                             const int32_t *__restrict__ nkeep, const int32_t *__restrict__ run_blk,
                             const int32_t *__restrict__ dmap, int U, int32_t *__restrict__ rcnt,
                             uint8_t *__restrict__ flag) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = u < U;
    // runs of dense blocks (dmap >= 0) go to the sparse-MFMA part instead
    const bool keep = valid && nkeep[u] > 0 && (dmap == nullptr || dmap[run_blk[u]] < 0);
    const int row = valid ? (int)(keys[rstart[u]] >> 32) : -1;
    if (valid) flag[u] = keep;
    const uint64_t kmask = __ballot(keep);
    const int prev = __shfl_up(row, 1, 64), next = __shfl_down(row, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || row != prev);
    const bool tail = lane == 63 || row != next;
    if (valid && tail) {
        const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
        const int start = 63 - __builtin_clzll(heads & upto);
        const int c = __popcll(kmask & upto & ~((1ull << start) - 1));
        if (c) atomicAdd(&rcnt[row], c);
    }
}

__global__ void k_iota_i(int32_t *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_scales(const uint32_t *__restrict__ absmax, int nb, int fp8, float *__restrict__ sc) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    float m = __uint_as_float(absmax[b]);
    sc[b] = (fp8 && m > 0.f) ? m / 448.0f : 1.0f;
}

__global__ void k_emit(const int32_t *__restrict__ sel_runs, int E, const int32_t *__restrict__ run_blk,
                       const uint64_t *__restrict__ keys, const int32_t *__restrict__ rstart,
                       const float *__restrict__ kv, const uint8_t *__restrict__ kidx,
                       const float *__restrict__ sc, const int32_t *__restrict__ blk_seg,
                       const uint16_t *__restrict__ blk_sel, int M, int dtype,
                       int32_t *__restrict__ ent_row, int32_t *__restrict__ ent_col,
                       int32_t *__restrict__ ent_blk, uint8_t *__restrict__ ent_idx,
                       uint8_t *__restrict__ ent_val) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int u = sel_runs[e];
    const int b = run_blk[u];
    ent_row[e] = (int32_t)(keys[rstart[u]] >> 32);
    // selector indices -> columns: first value absolute, second as offset
    const uint32_t code = kidx[u], sel = blk_sel[b];
    ent_col[e] = blk_seg[b] * M + (int)((sel >> (4 * (code & 3u))) & 15u);
    ent_idx[e] = (code & 0x10u) ? (uint8_t)(0x10u | ((sel >> (4 * ((code >> 2) & 3u))) & 15u)) : (uint8_t)0;
    if (dtype == VNM_DT_FP8) ent_blk[e] = b;
    const float v0 = kv[2 * u], v1 = kv[2 * u + 1];
    if (dtype == VNM_DT_F32) {
        float *p = (float *)ent_val;
        p[2 * e] = v0;
        p[2 * e + 1] = v1;
    } else if (dtype == VNM_DT_BF16) {
        uint16_t *p = (uint16_t *)ent_val;
        p[2 * e] = f32_to_bf16(v0);
        p[2 * e + 1] = f32_to_bf16(v1);
    } else if (dtype == VNM_DT_F16) {
        __half *p = (__half *)ent_val;
        p[2 * e] = __float2half(v0);
        p[2 * e + 1] = __float2half(v1);
    } else {
        const float s = sc[b];
        ent_val[2 * e] = f32_to_e4m3(v0 / s);
        ent_val[2 * e + 1] = f32_to_e4m3(v1 / s);
    }
}

// ------------------------------------------------------------ SpMM

// ------------------------------------------------------------ value stream

// Values per row entry (1 or 2), and entry E's count 0 (the scan's total).
__global__ void k_val_counts(const uint8_t *__restrict__ ent_idx, int E, int32_t *__restrict__ nv) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) nv[e] = (ent_idx[e] & 0x10u) ? 2 : 1;
    else if (e == E) nv[e] = 0;
}

template <int DT>
__device__ __forceinline__ void emit_values(int e, int o, int M, const int32_t *__restrict__ ent_row,
                                            const int32_t *__restrict__ ent_col, const uint8_t *__restrict__ ent_idx,
                                            const uint8_t *__restrict__ ent_val, const int32_t *__restrict__ ent_blk,
                                            const float *__restrict__ blk_scale, int32_t *__restrict__ vrow,
                                            int32_t *__restrict__ vcol, float *__restrict__ vw) {
    const int r = ent_row[e], c0 = ent_col[e];
    const uint32_t code = ent_idx[e];
    const float sc = DT == VNM_DT_FP8 ? blk_scale[ent_blk[e]] : 1.f;
    vrow[o] = r;
    vcol[o] = c0;
    vw[o] = load_aval<DT>(ent_val, 2 * e) * sc;
    if (code & 0x10u) {
        vrow[o + 1] = r;
        vcol[o + 1] = (c0 & ~(M - 1)) | (int)(code & 15u);
        vw[o + 1] = load_aval<DT>(ent_val, 2 * e + 1) * sc;
    }
}

// The SpMM's value stream: row entry e's values (its first, then its second)
// at voff[e] as (row, B row, weight), the weight being the stored value times
// the block scale (FP8; 1 otherwise) rounded to FP32 once.
__global__ void k_values(int E, int M, int dtype, const int32_t *__restrict__ voff,
                         const int32_t *__restrict__ ent_row, const int32_t *__restrict__ ent_col,
                         const uint8_t *__restrict__ ent_idx, const uint8_t *__restrict__ ent_val,
                         const int32_t *__restrict__ ent_blk, const float *__restrict__ blk_scale,
                         int32_t *__restrict__ vrow, int32_t *__restrict__ vcol, float *__restrict__ vw) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int o = voff[e];
    switch (dtype) {
        case VNM_DT_F32:
            emit_values<VNM_DT_F32>(e, o, M, ent_row, ent_col, ent_idx, ent_val, ent_blk, blk_scale, vrow, vcol, vw);
            break;
        case VNM_DT_BF16:
            emit_values<VNM_DT_BF16>(e, o, M, ent_row, ent_col, ent_idx, ent_val, ent_blk, blk_scale, vrow, vcol, vw);
            break;
        case VNM_DT_F16:
            emit_values<VNM_DT_F16>(e, o, M, ent_row, ent_col, ent_idx, ent_val, ent_blk, blk_scale, vrow, vcol, vw);
            break;
        default:
            emit_values<VNM_DT_FP8>(e, o, M, ent_row, ent_col, ent_idx, ent_val, ent_blk, blk_scale, vrow, vcol, vw);
            break;
    }
}

// Row pointers over values: vrp[r] = voff[rowptr[r]], r <= R.
__global__ void k_vrowptr(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ voff, int R,
                          int32_t *__restrict__ vrp) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= R) vrp[r] = voff[rowptr[r]];
}

// ------------------------------------------------------------ SpMM kernel

template <int BT, int VEC>
__global__ __launch_bounds__(256) void k_spmm(SpmmArgs a, int T, int items, int per_g) {
    const int wv = threadIdx.x >> 6;
    // workgroups are dealt round-robin over the 8 XCDs
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    int tile, iwg;
    if (8 % T == 0) {  // tile t on XCDs [t*8/T, (t+1)*8/T)
        const int gs = 8 / T;
        tile = xcd / gs;
        iwg = (xcd % gs) * per_g + slot;
    } else {
        const int lwg = xcd * per_g + slot;
        tile = lwg % T;
        iwg = lwg / T;
    }
    // wave-uniform; readfirstlane lets the entry reads be scalar loads
    const int item = __builtin_amdgcn_readfirstlane(iwg * 4 + wv);
    if (item >= items) return;
    spmm_unit<BT, VEC>(a, item, __builtin_amdgcn_readfirstlane(tile), T);
}

// Fix-up pass: rows cut into chunks get the sum of their partial-sum slots
// in item order (one wave each); a matrix without work items, the bias or
// zero in every row.
__global__ __launch_bounds__(256) void k_spmm_fix(const int4 *__restrict__ fix, int nfix, const float *__restrict__ part,
                                                  int pn, int n, float *__restrict__ C, int ldc,
                                                  const float *__restrict__ bias, const uint8_t *__restrict__ dvb,
                                                  int V) {
    // wave-uniform, so the list entry is a scalar load
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= nfix) return;
    const int4 f = fix[w];
    fix_row(C + (size_t)f.x * ldc, part, pn, f.y, f.z, n, bias, dvb != nullptr && dvb[f.x / V] != 0);
}

// ------------------------------------------------------------ host

static std::vector<uint32_t> subsets4(int M) {
    std::vector<uint32_t> out;
    if (M <= 4) {
        out.push_back((1u << M) - 1);
        return out;
    }
    for (int a = 0; a < M; ++a)
        for (int b = a + 1; b < M; ++b)
            for (int c = b + 1; c < M; ++c)
                for (int d = c + 1; d < M; ++d)
                    out.push_back((1u << (M - 1 - a)) | (1u << (M - 1 - b)) | (1u << (M - 1 - c)) |
                                  (1u << (M - 1 - d)));
    return out;
}

// SpMM work items from the value row pointers (host pass, once per packed matrix):
// consecutive whole rows while they fit in ITEM values; a row longer than
// ITEM starts a new item and is cut into ITEM-value chunks, each with a
// partial-sum slot of its own (consecutive slots), the last chunk staying
// open for the rows that follow.  So an item's only cut row is its first.
// The fix-up list holds the cut rows with their slot range and the rows of
// runs of more than GAP_MAX rows that keep nothing (the SpMM's item walk
// writes the shorter runs; without any item, the list holds every row).
static int build_items(vnm_packed *p, const int32_t *d_vrp) {
    hipStream_t st = p->ctx->stream;
    const int R = p->nrows;
    std::vector<int32_t> rp(R + 1);
    VNM_HIP_TRY(hipMemcpyAsync(rp.data(), d_vrp, (size_t)(R + 1) * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    std::vector<int32_t> beg, slot;
    std::vector<int4> fix;
    int cur_len = 0, nslots = 0;
    bool open = false;
    for (int r = 0; r < R; ++r) {
        const int b = rp[r], L = rp[r + 1] - b;
        if (L == 0) continue;
        if (L <= ITEM) {
            if (open && cur_len + L > ITEM) open = false;
            if (!open) {
                beg.push_back(b);
                slot.push_back(-1);
                open = true;
                cur_len = 0;
            }
            cur_len += L;
        } else {
            const int nch = (L + ITEM - 1) / ITEM;
            fix.push_back(make_int4(r, nslots, nch, 0));
            for (int c = 0; c < nch; ++c) {
                beg.push_back(b + c * ITEM);
                slot.push_back(nslots++);
            }
            open = true;
            cur_len = L - (nch - 1) * ITEM;
        }
    }
    beg.push_back(rp[R]);
    if (slot.empty()) {
        for (int r = 0; r < R; ++r) fix.push_back(make_int4(r, 0, 0, 0));
    } else {
        // runs of more than GAP_MAX rows without entries: one fix-up wave per row
        for (int r = 0; r < R;) {
            if (rp[r + 1] != rp[r]) { ++r; continue; }
            int e = r;
            while (e < R && rp[e + 1] == rp[e]) ++e;
            if (e - r > GAP_MAX)
                for (int q = r; q < e; ++q) fix.push_back(make_int4(q, 0, 0, 0));
            r = e;
        }
    }
    p->nitems = (int)slot.size();
    p->nslots = nslots;
    p->nfix = (int)fix.size();
    int rc;
    if ((rc = p->item_beg.alloc(beg.size())) || (rc = p->item_slot.alloc(std::max<size_t>(slot.size(), 1))) ||
        (rc = p->fix.alloc(std::max<size_t>(fix.size(), 1))))
        return rc;
    VNM_HIP_TRY(hipMemcpyAsync(p->item_beg.p, beg.data(), beg.size() * 4, hipMemcpyHostToDevice, st));
    if (!slot.empty())
        VNM_HIP_TRY(hipMemcpyAsync(p->item_slot.p, slot.data(), slot.size() * 4, hipMemcpyHostToDevice, st));
    if (!fix.empty())
        VNM_HIP_TRY(hipMemcpyAsync(p->fix.p, fix.data(), fix.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int pack(vnm_graph *g, int V, int dtype, int dense_min, vnm_packed **out, int64_t *retained_only) {
    const int M = g->M;
    if (dense_min < 0 || (dense_min > 0 && (dtype != VNM_DT_BF16 || V % 16 != 0)))
        return set_error(VNM_E_ARG, "the sparse-MFMA part needs BF16 values and V a multiple of 16 "
                                    "(dense_min %d, dtype %d, V %d)", dense_min, dtype, V);
    if (V <= 0 || V > 256) return set_error(VNM_E_ARG, "V must be in [1, 256] (got %d)", V);
    if (M != 4 && M != 8 && M != 16) return set_error(VNM_E_ARG, "M must be 4, 8 or 16");
    if (dtype < VNM_DT_F32 || dtype > VNM_DT_FP8) return set_error(VNM_E_ARG, "bad dtype %d", dtype);
    if (g->col_overflow)
        return set_error(VNM_E_ARG, "entries lie past the padded column count after a relabel of a matrix "
                                    "with fewer columns than rows");
    hipStream_t st = g->ctx->stream;
    const int R = g->nrows;
    const int nseg = (g->ncols + M - 1) / M;
    const int nblk = (R + V - 1) / V;
    const int64_t nnz = g->nnz;
    const size_t e1 = std::max<int64_t>(nnz, 1);
    int rc;
    DBuf<uint64_t> keys, keys2, bkey, bkey2;
    DBuf<int32_t> idx, idx2, head, excl, rstart, rlen, rid, rid2, bhead, bexcl, blk_first, run_blk,
        nkeep, rcnt, sel_runs, cnt;
    DBuf<uint32_t> bits, blk_mask, absmax, dsubs;
    DBuf<float> kv;
    DBuf<uint8_t> kidx, flag;
    DBuf<unsigned long long> dret;
    DBuf<int32_t> longb, nlong;
    if ((rc = keys.alloc(e1)) || (rc = keys2.alloc(e1)) || (rc = idx.alloc(e1)) ||
        (rc = idx2.alloc(e1)) || (rc = head.alloc(e1)) || (rc = excl.alloc(e1)) ||
        (rc = dret.alloc(1)) || (rc = cnt.alloc(4)))
        return rc;
    const int gb = grid_of(nnz, 256);
    hipLaunchKernelGGL(k_nz_keys, dim3(gb), dim3(256), 0, st, g->row.p, g->col.p, g->val.p, nnz,
                       nseg * M, keys.p, idx.p);
    if ((rc = sort_pairs_u64_i32(keys.p, keys2.p, idx.p, idx2.p, nnz, 64, st))) return rc;
    // count of nonzero entries = position of the first ~0 key
    std::vector<uint64_t> hk(1);
    int64_t nz = nnz;
    {
        // binary search on the device array through small copies
        int64_t lo = 0, hi = nnz;
        while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            VNM_HIP_TRY(hipMemcpyAsync(hk.data(), keys2.p + mid, 8, hipMemcpyDeviceToHost, st));
            VNM_HIP_TRY(hipStreamSynchronize(st));
            if (hk[0] == ~0ull) hi = mid; else lo = mid + 1;
        }
        nz = lo;
    }
    int U = 0;
    if (nz > 0) {
        hipLaunchKernelGGL(k_pk_heads, dim3(grid_of(nz, 256)), dim3(256), 0, st, keys2.p, nz, M, head.p);
        if ((rc = scan_excl_i32(head.p, excl.p, nz, st))) return rc;
        int32_t last[2];
        VNM_HIP_TRY(hipMemcpyAsync(&last[0], excl.p + nz - 1, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(&last[1], head.p + nz - 1, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        U = last[0] + last[1];
    }
    const size_t u1 = std::max(U, 1);
    if ((rc = bits.alloc(u1)) || (rc = rstart.alloc(u1)) || (rc = rlen.alloc(u1)) ||
        (rc = bkey.alloc(u1)) || (rc = bkey2.alloc(u1)) || (rc = rid.alloc(u1)) ||
        (rc = rid2.alloc(u1)) || (rc = bhead.alloc(u1)) || (rc = bexcl.alloc(u1)) ||
        (rc = run_blk.alloc(u1)) || (rc = nkeep.alloc(u1)) || (rc = kv.alloc(2 * u1)) ||
        (rc = kidx.alloc(u1)) || (rc = flag.alloc(u1)) || (rc = sel_runs.alloc(u1)) ||
        (rc = rcnt.alloc(R + 1)))
        return rc;
    int nb = 0;
    std::vector<uint32_t> subs = subsets4(M);
    if ((rc = dsubs.alloc(subs.size()))) return rc;
    VNM_HIP_TRY(hipMemcpyAsync(dsubs.p, subs.data(), subs.size() * 4, hipMemcpyHostToDevice, st));
    VNM_HIP_TRY(hipMemsetAsync(dret.p, 0, 8, st));
    if (U > 0) {
        VNM_HIP_TRY(hipMemsetAsync(bits.p, 0, (size_t)U * 4, st));
        VNM_HIP_TRY(hipMemsetAsync(rlen.p, 0, (size_t)U * 4, st));
        hipLaunchKernelGGL(k_pk_runs, dim3(grid_of(nz, 256)), dim3(256), 0, st, keys2.p, head.p,
                           excl.p, nz, M, V, nblk, bits.p, rstart.p, rlen.p, bkey.p, rid.p);
        const int kb = 32 + bits_for((int64_t)nseg * nblk);
        if ((rc = sort_pairs_u64_i32(bkey.p, bkey2.p, rid.p, rid2.p, U, std::min(64, kb), st))) return rc;
        hipLaunchKernelGGL(k_blk_heads, dim3(grid_of(U, 256)), dim3(256), 0, st, bkey2.p, U, bhead.p);
        if ((rc = scan_excl_i32(bhead.p, bexcl.p, U, st))) return rc;
        int32_t last[2];
        VNM_HIP_TRY(hipMemcpyAsync(&last[0], bexcl.p + U - 1, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(&last[1], bhead.p + U - 1, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        nb = last[0] + last[1];
    }
    vnm_packed *p = new vnm_packed();
    p->ctx = g->ctx;
    p->V = V; p->M = M; p->nrows = R; p->ncols = g->ncols; p->dtype = dtype; p->nblocks = nb;
    const size_t b1 = std::max(nb, 1);
    if ((rc = blk_first.alloc(b1)) || (rc = blk_mask.alloc(b1)) || (rc = absmax.alloc(b1)) ||
        (rc = p->blk_seg.alloc(b1)) || (rc = p->blk_sel.alloc(b1)) || (rc = p->blk_scale.alloc(b1))) {
        delete p;
        return rc;
    }
    if (nb > 0) {
        hipLaunchKernelGGL(k_blk_first, dim3(grid_of(U, 256)), dim3(256), 0, st, bhead.p, bexcl.p, U,
                           blk_first.p);
        // short blocks one per lane; the rest one per wave, from a list the
        // lane kernel fills (its length read on the device: no host wait)
        if ((rc = longb.alloc(b1)) || (rc = nlong.alloc(1))) {
            delete p;
            return rc;
        }
        VNM_HIP_TRY(hipMemsetAsync(nlong.p, 0, 4, st));
        hipLaunchKernelGGL(k_blk_select_lane, dim3(std::min(grid_of(nb, 256), 2048)), dim3(256), 0, st, bkey2.p, rid2.p, U, nb,
                           M, nblk, bits.p, dsubs.p, (int)subs.size(), blk_first.p, p->blk_seg.p, p->blk_sel.p,
                           blk_mask.p, run_blk.p, dret.p, longb.p, nlong.p);
        hipLaunchKernelGGL(k_blk_select, dim3(std::min((nb + 3) / 4, 4096)), dim3(256), 0, st, bkey2.p, rid2.p,
                           U, nb, M, nblk, bits.p, dsubs.p, (int)subs.size(), blk_first.p, p->blk_seg.p,
                           p->blk_sel.p, blk_mask.p, run_blk.p, dret.p, longb.p, nlong.p);
        VNM_HIP_TRY(hipMemsetAsync(absmax.p, 0, (size_t)nb * 4, st));
        hipLaunchKernelGGL(k_run_keep, dim3(grid_of(U, 256)), dim3(256), 0, st, keys2.p, idx2.p,
                           g->val.p, U, M, rstart.p, rlen.p, run_blk.p, blk_mask.p, kv.p, kidx.p,
                           nkeep.p, absmax.p);
        hipLaunchKernelGGL(k_scales, dim3(grid_of(nb, 256)), dim3(256), 0, st, absmax.p, nb,
                           dtype == VNM_DT_FP8, p->blk_scale.p);
    }
    unsigned long long ret = 0;
    VNM_HIP_TRY(hipMemcpyAsync(&ret, dret.p, 8, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    p->retained = (int64_t)ret;
    VNM_LAUNCH_CHECK("pack select");
    if (retained_only) {
        *retained_only = p->retained;
        delete p;
        return VNM_OK;
    }
    // dense blocks to the sparse-MFMA part
    DBuf<int32_t> dmap;
    if (dense_min > 0 && nb > 0 &&
        (rc = build_dense(p, dense_min, U, nb, nblk, run_blk.p, nkeep.p, blk_first.p, bkey2.p, keys2.p, rstart.p,
                          kv.p, kidx.p, dmap))) {
        delete p;
        return rc;
    }
    // row entries (runs of the other blocks that keep something), row-major
    VNM_HIP_TRY(hipMemsetAsync(rcnt.p, 0, (size_t)(R + 1) * 4, st));
    int E = 0;
    if (U > 0) {
        hipLaunchKernelGGL(k_row_counts, dim3(grid_of(U, 256)), dim3(256), 0, st, keys2.p, rstart.p,
                           nkeep.p, run_blk.p, p->ngroups > 0 ? dmap.p : nullptr, U, rcnt.p, flag.p);
        hipLaunchKernelGGL(k_iota_i, dim3(grid_of(U, 256)), dim3(256), 0, st, rid.p, U);
        if ((rc = select_flagged_i32(rid.p, flag.p, sel_runs.p, cnt.p, U, st))) { delete p; return rc; }
        VNM_HIP_TRY(hipMemcpyAsync(&E, cnt.p, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
    }
    p->nent = E;
    const int vs = dtype == VNM_DT_F32 ? 4 : dtype == VNM_DT_FP8 ? 1 : 2;
    if ((rc = p->rowptr.alloc(R + 1)) || (rc = p->ent_col.alloc(std::max(E, 1))) ||
        (rc = p->ent_blk.alloc(dtype == VNM_DT_FP8 ? std::max(E, 1) : 1)) ||
        (rc = p->ent_row.alloc(std::max(E, 1))) ||
        (rc = p->ent_idx.alloc(std::max(E, 1))) ||
        (rc = p->ent_val.alloc((size_t)std::max(E, 1) * 2 * vs))) {
        delete p;
        return rc;
    }
    if ((rc = scan_excl_i32(rcnt.p, p->rowptr.p, R + 1, st))) { delete p; return rc; }
    if (E > 0)
        hipLaunchKernelGGL(k_emit, dim3(grid_of(E, 256)), dim3(256), 0, st, sel_runs.p, E, run_blk.p,
                           keys2.p, rstart.p, kv.p, kidx.p, p->blk_scale.p, p->blk_seg.p, p->blk_sel.p,
                           M, dtype, p->ent_row.p, p->ent_col.p, p->ent_blk.p, p->ent_idx.p, p->ent_val.p);
    g->ctx->launches += 17;
    VNM_LAUNCH_CHECK("pack emit");
    // the value stream and its row pointers
    DBuf<int32_t> nv, voff, vrp;
    if ((rc = nv.alloc((size_t)E + 1)) || (rc = voff.alloc((size_t)E + 1)) || (rc = vrp.alloc((size_t)R + 1))) {
        delete p;
        return rc;
    }
    hipLaunchKernelGGL(k_val_counts, dim3(grid_of((int64_t)E + 1, 256)), dim3(256), 0, st, p->ent_idx.p, E, nv.p);
    if ((rc = scan_excl_i32(nv.p, voff.p, (int64_t)E + 1, st))) { delete p; return rc; }
    int NV = 0;
    VNM_HIP_TRY(hipMemcpyAsync(&NV, voff.p + E, 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    p->nval = NV;
    // VTAIL zero records past the end: the SpMM reads B rows and weights a
    // block of values at a time
    const size_t vn = (size_t)NV + VTAIL;
    if ((rc = p->vrow.alloc(vn)) || (rc = p->vcol.alloc(vn)) || (rc = p->vw.alloc(vn))) {
        delete p;
        return rc;
    }
    VNM_HIP_TRY(hipMemsetAsync(p->vrow.p + NV, 0, VTAIL * 4, st));
    VNM_HIP_TRY(hipMemsetAsync(p->vcol.p + NV, 0, VTAIL * 4, st));
    VNM_HIP_TRY(hipMemsetAsync(p->vw.p + NV, 0, VTAIL * 4, st));
    if (E > 0)
        hipLaunchKernelGGL(k_values, dim3(grid_of(E, 256)), dim3(256), 0, st, E, M, dtype, voff.p, p->ent_row.p,
                           p->ent_col.p, p->ent_idx.p, p->ent_val.p, p->ent_blk.p, p->blk_scale.p,
                           p->vrow.p, p->vcol.p, p->vw.p);
    hipLaunchKernelGGL(k_vrowptr, dim3(grid_of((int64_t)R + 1, 256)), dim3(256), 0, st, p->rowptr.p, voff.p, R,
                       vrp.p);
    g->ctx->launches += 4;
    VNM_LAUNCH_CHECK("pack values");
    if ((rc = build_items(p, vrp.p))) { delete p; return rc; }
    *out = p;
    return VNM_OK;
}

// Rows of n floats of SpMM scratch: the partial-sum slots of rows cut into
// chunks, then the dense units' slots (V rows each).
size_t scratch_rows(const vnm_packed *p) { return (size_t)p->nslots + (size_t)p->ndslots * p->V; }

// Partial-sum slots for n output columns, grown on demand.
int ensure_part(vnm_packed *p, int n) {
    const size_t need = std::max<size_t>(scratch_rows(p) * n, 1);
    if (p->part.n >= need) return VNM_OK;
    VNM_HIP_TRY(hipStreamSynchronize(p->ctx->stream));  // the old buffer may be in use
    return p->part.alloc(need);
}

int bsize_of(int bt) { return bt == VNM_DT_F32 ? 4 : bt == VNM_DT_BF16 ? 2 : 1; }

// Values per lane usable for (bt, n): 2 to 16 bytes per lane load, and the
// 64 * vec column tile divides n.
bool vec_ok(int bt, int n, int vec) {
    if (vec != 1 && vec != 2 && vec != 4 && vec != 8 && vec != 16) return false;
    const int bytes = vec * bsize_of(bt);
    return bytes >= 2 && bytes <= 16 && n % (64 * vec) == 0;
}

SpmmArgs packed_args(const vnm_packed *p, const void *B, int ldb, const float *bs, float *C, int ldc,
                     const float *bias, float *part, int pn) {
    SpmmArgs a;
    a.item_beg = p->item_beg.p;
    a.item_slot = p->item_slot.p;
    a.vrow = p->vrow.p;
    a.vcol = p->vcol.p;
    a.vw = p->vw.p;
    a.B = (const uint8_t *)B;
    a.b_scales = bs;
    a.C = C;
    a.part = part;
    a.bias = bias;
    a.items = p->nitems;
    a.ldb = ldb;
    a.ldc = ldc;
    a.pn = pn;
    a.b_rows = p->ncols;
    a.c_rows = p->nrows;
    a.dvb = p->ngroups > 0 ? p->dvb.p : nullptr;
    a.V = p->V;
    return a;
}

template <int BT, int VEC>
static void launch_v(const SpmmArgs &args, int n, hipStream_t st) {
    const int T = n / (64 * VEC);
    const int items = args.items;
    const int iwg = (items + 3) / 4;
    int per_g;
    if (8 % T == 0) {
        const int gs = 8 / T;
        per_g = (iwg + gs - 1) / gs;
    } else {
        per_g = (iwg * T + 7) / 8;
    }
    hipLaunchKernelGGL((k_spmm<BT, VEC>), dim3(8 * per_g), dim3(256), 0, st, args, T, items, per_g);
}

// The value stream carries FP32 weights, so the kernel depends on B's dtype
// and the values per lane only.
static void launch_b(const SpmmArgs &args, int vec, int n, int bt, hipStream_t st) {
    if (bt == VNM_DT_F32) {
        if (vec == 4) launch_v<VNM_DT_F32, 4>(args, n, st);
        else if (vec == 2) launch_v<VNM_DT_F32, 2>(args, n, st);
        else launch_v<VNM_DT_F32, 1>(args, n, st);
    } else if (bt == VNM_DT_BF16) {
        if (vec == 8) launch_v<VNM_DT_BF16, 8>(args, n, st);
        else if (vec == 4) launch_v<VNM_DT_BF16, 4>(args, n, st);
        else if (vec == 2) launch_v<VNM_DT_BF16, 2>(args, n, st);
        else launch_v<VNM_DT_BF16, 1>(args, n, st);
    } else {
        if (vec == 16) launch_v<VNM_DT_FP8, 16>(args, n, st);
        else if (vec == 8) launch_v<VNM_DT_FP8, 8>(args, n, st);
        else if (vec == 4) launch_v<VNM_DT_FP8, 4>(args, n, st);
        else launch_v<VNM_DT_FP8, 2>(args, n, st);
    }
}

// One SpMM with `vec` values per lane and the partial-sum slots at `part`
// (nslots x n floats): the gather kernel, then the fix-up pass.
int spmm_run(vnm_packed *p, int vec, const void *B, int ldb, int n, int bt, const float *bs, float *C,
                    int ldc, const float *bias, float *part, bool timed) {
    hipStream_t st = p->ctx->stream;
    const SpmmArgs args = packed_args(p, B, ldb, bs, C, ldc, bias, part, n);
    // the dense blocks' 2:4 products first: the row-entry pass adds to them
    int rc = spmm_dense(p, (const uint16_t *)B, ldb, n, C, ldc, part + (size_t)p->nslots * n, st, timed);
    if (rc) return rc;
    if (p->nitems > 0) {
        KernelTimer kt(timed ? p->ctx : nullptr, "k_spmm");
        launch_b(args, vec, n, bt, st);
        kt.stop();
        p->ctx->launches++;
    }
    if (p->nfix > 0) {
        hipLaunchKernelGGL(k_spmm_fix, dim3((p->nfix + 3) / 4), dim3(256), 0, st, p->fix.p, p->nfix, part, n, n, C,
                           ldc, bias, args.dvb, p->V);
        p->ctx->launches++;
    }
    VNM_LAUNCH_CHECK("k_spmm");
    return VNM_OK;
}

// spmm_run with the packed matrix's own slots.
static int spmm_v(vnm_packed *p, int vec, const void *B, int ldb, int n, int bt, const float *bs, float *C,
                  int ldc, const float *bias, bool timed) {
    int rc = ensure_part(p, n);
    if (rc) return rc;
    return spmm_run(p, vec, B, ldb, n, bt, bs, C, ldc, bias, p->part.p, timed);
}

int shape_key(int bt, int n) { return (bt << 24) | n; }

// Autotuner: time every usable values-per-lane setting on this matrix and
// shape (one warm-up and three timed calls each, HIP events on the context
// stream) and keep the fastest; the choice stays with the packed matrix.
static int tune_vec(vnm_packed *p, const void *B, int ldb, int n, int bt, const float *bs, float *C, int ldc,
                    const float *bias, int *best_v) {
    hipStream_t st = p->ctx->stream;
    hipEvent_t e0, e1;
    VNM_HIP_TRY(hipEventCreate(&e0));
    VNM_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int bv = 0, rc = VNM_OK;
    for (int v = 16; v >= 1 && rc == VNM_OK; v /= 2) {
        if (!vec_ok(bt, n, v)) continue;
        if ((rc = spmm_v(p, v, B, ldb, n, bt, bs, C, ldc, bias, false))) break;
        if (hipEventRecord(e0, st) != hipSuccess) { rc = set_error(VNM_E_HIP, "event record failed"); break; }
        for (int i = 0; i < 3 && rc == VNM_OK; ++i) rc = spmm_v(p, v, B, ldb, n, bt, bs, C, ldc, bias, false);
        if (rc) break;
        float ms = 0.f;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = set_error(VNM_E_HIP, "event timing failed");
            break;
        }
        if (getenv("VNM_DEBUG")) fprintf(stderr, "[vnm] spmm tune bt=%d n=%d vec=%d %.4f ms\n", bt, n, v, ms / 3);
        if (ms < best) { best = ms; bv = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best_v = bv;
    return VNM_OK;
}

int spmm_check(const vnm_packed *p, const void *B, int ldb, int n, int bt, const float *bs,
                      const float *C, int ldc) {
    if (!p || !B || !C || n <= 0 || n % 64 != 0 || ldb < n || ldc < n)
        return set_error(VNM_E_ARG, "spmm: n must be a positive multiple of 64 with ldb, ldc >= n");
    if (bt != VNM_DT_F32 && bt != VNM_DT_BF16 && bt != VNM_DT_FP8)
        return set_error(VNM_E_ARG, "spmm: B dtype must be F32, BF16 or FP8");
    if (bt == VNM_DT_FP8 && !bs) return set_error(VNM_E_ARG, "spmm: FP8 B needs per-128-row scales");
    // vector loads of up to 16 bytes of B, stores of up to 16 bytes of C
    if ((ldb * bsize_of(bt)) % 16 != 0 || (ldc % 4) != 0 || ((uintptr_t)B & 15) || ((uintptr_t)C & 15))
        return set_error(VNM_E_ARG, "spmm: B and C must be 16-byte aligned with rows of whole 16-byte "
                                    "units (ldb * sizeof(B value) and 4 * ldc multiples of 16)");
    if ((uint64_t)p->ncols * (uint64_t)ldb * (uint64_t)bsize_of(bt) > 0xFFFFFFFFull)
        return set_error(VNM_E_ARG, "spmm: B (ncols x ldb) must be smaller than 4 GiB");
    if (p->ngroups > 0 && bt != VNM_DT_BF16)
        return set_error(VNM_E_ARG, "spmm: a matrix packed with a sparse-MFMA part takes BF16 B");
    return VNM_OK;
}

int spmm(vnm_packed *p, const void *B, int ldb, int n, int bt, const float *bs, float *C, int ldc,
         const float *bias) {
    int rc = spmm_check(p, B, ldb, n, bt, bs, C, ldc);
    if (rc) return rc;
    int v = 0;
    auto it = p->shapes.find(shape_key(bt, n));
    if (it != p->shapes.end()) {
        v = it->second;
    } else {
        if ((rc = tune_vec(p, B, ldb, n, bt, bs, C, ldc, bias, &v))) return rc;
        p->shapes[shape_key(bt, n)] = v;
    }
    return spmm_v(p, v, B, ldb, n, bt, bs, C, ldc, bias, true);
}

}  // namespace vnm

using namespace vnm;

extern "C" {

int vnm_retained_nnz(vnm_graph *g, int V, int N, int64_t *out) {
    if (!g || !out) return set_error(VNM_E_ARG, "null argument");
    if (N != 2) return set_error(VNM_E_ARG, "the V:N:M format keeps N = 2 per row (got %d)", N);
    return pack(g, V, VNM_DT_F32, 0, nullptr, out);
}

int vnm_pack(vnm_graph *g, int V, int dtype, vnm_packed **out) {
    if (!g || !out) return set_error(VNM_E_ARG, "null argument");
    return pack(g, V, dtype, 0, out, nullptr);
}

int vnm_pack_ex(vnm_graph *g, int V, int dtype, int dense_min, vnm_packed **out) {
    if (!g || !out) return set_error(VNM_E_ARG, "null argument");
    return pack(g, V, dtype, dense_min, out, nullptr);
}

int vnm_packed_info_get(const vnm_packed *p, vnm_packed_info *info) {
    if (!p || !info) return set_error(VNM_E_ARG, "null argument");
    info->V = p->V;
    info->M = p->M;
    info->nrows = p->nrows;
    info->ncols = p->ncols;
    info->nblocks = p->nblocks;
    info->dtype = p->dtype;
    info->retained_nnz = p->retained;
    const int vs = p->dtype == VNM_DT_F32 ? 4 : p->dtype == VNM_DT_FP8 ? 1 : 2;
    const int blk = p->dtype == VNM_DT_FP8 ? 4 : 0;  // per-entry block index for the FP8 scale
    info->bytes = 4ll * (p->nrows + 1) + (int64_t)p->nent * (4 + 4 + 1 + 2 * vs + blk) +
                  (int64_t)p->nblocks * (4 + 2 + 4);
    // the SpMM's value stream: (row, B row, FP32 weight) per value
    info->bytes += (int64_t)p->nval * 12;
    // what k_spmm reads: the value stream and the item table (first value
    // and slot of each item)
    info->spmm_bytes = (int64_t)p->nval * 12 + (int64_t)p->nitems * 8;
    info->items = p->nitems;
    info->slots = (int32_t)scratch_rows(p);
    info->dense_blocks = p->ndense;
    info->dense_chunks = p->nchunks;
    // the sparse-MFMA part reads per chunk its 64 column indices and per
    // 16-row slice 64 lanes x (16 value bytes + 4 position bytes)
    info->spmm_bytes += (int64_t)p->nchunks * (64 * 4 + (int64_t)(p->V / 16) * 64 * 20) + (int64_t)p->ngroups * 16;
    info->bytes += (int64_t)p->nchunks * (64 * 4 + (int64_t)(p->V / 16) * 64 * 20) + (int64_t)p->ngroups * 16 +
                   (p->ngroups ? (int64_t)((p->nrows + p->V - 1) / p->V) : 0);
    return VNM_OK;
}

void vnm_packed_destroy(vnm_packed *p) {
    if (!p) return;
    if (p->ctx && p->ctx->stream) (void)hipStreamSynchronize(p->ctx->stream);
    delete p;
}

int vnm_spmm(vnm_packed *p, const void *B, int ldb, int n, int b_dtype, const float *b_scales,
             float *C, int ldc) {
    return spmm(p, B, ldb, n, b_dtype, b_scales, C, ldc, nullptr);
}

int vnm_spmm_set_shape(vnm_packed *p, int b_dtype, int n, int cols_per_lane) {
    if (!p) return set_error(VNM_E_ARG, "null argument");
    if (cols_per_lane == 0) {
        p->shapes.erase(shape_key(b_dtype, n));
        return VNM_OK;
    }
    if (!vec_ok(b_dtype, n, cols_per_lane))
        return set_error(VNM_E_ARG, "spmm: %d columns per lane do not fit n = %d for B dtype %d",
                         cols_per_lane, n, b_dtype);
    p->shapes[shape_key(b_dtype, n)] = cols_per_lane;
    return VNM_OK;
}

int vnm_spmm_get_shape(const vnm_packed *p, int b_dtype, int n, int32_t *cols_per_lane) {
    if (!p || !cols_per_lane) return set_error(VNM_E_ARG, "null argument");
    auto it = p->shapes.find(shape_key(b_dtype, n));
    *cols_per_lane = it == p->shapes.end() ? 0 : it->second;
    return VNM_OK;
}

}  // extern "C"
