// The packed V:N:M format and the SpMM device code shared by the single and
// batched SpMM kernels (pack.hip, batch.hip).  Format and kernel design are
// described in pack.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstring>
#include <type_traits>
#include <map>
#include <memory>

#include "graph.h"
#include "hip_common.h"

struct vnm_packed {
    vnm_ctx *ctx = nullptr;
    int32_t V = 0, M = 0, nrows = 0, ncols = 0, dtype = 0;
    int32_t nblocks = 0, nent = 0;
    int64_t retained = 0;
    vnm::DBuf<int32_t> rowptr, ent_row, ent_col, ent_blk, blk_seg;
    vnm::DBuf<uint8_t> ent_idx, ent_val;                   // values: 2 per entry, dtype-sized
    vnm::DBuf<uint16_t> blk_sel;
    vnm::DBuf<float> blk_scale;
    // SpMM work items: whole rows of at most ITEM row entries, and rows longer
    // than ITEM cut into ITEM-entry chunks whose partial sums go to slots
    int32_t nitems = 0, nslots = 0, nfix = 0;
    vnm::DBuf<int32_t> item_beg;                           // nitems + 1 entry offsets
    vnm::DBuf<int32_t> item_slot;                          // slot of the item's first row, or -1
    vnm::DBuf<int4> fix;                                   // {row, first slot, slots, 0}: rows the
                                                           // fix-up pass writes (empty or cut)
    vnm::DBuf<float> part;                                 // nslots x n partial sums
    // 2:4 sparse-MFMA part (dense.hip): blocks keeping >= dense_min values
    int32_t dense_min = 0, ndense = 0, nchunks = 0, ngroups = 0, nmulti = 0, ndslots = 0;
    vnm::DBuf<int4> dgrp;                                  // units {V-row block, first chunk, chunks, slot or -1}
    vnm::DBuf<int4> dmulti;                                // {V-row block, first slot, slots, 0}
    vnm::DBuf<int32_t> dcol;                               // chunks x 64 B rows (-1: none)
    vnm::DBuf<uint32_t> dfrag;                             // chunks x V/16 x 64 lanes x 8 BF16
    vnm::DBuf<uint32_t> didx;                              // chunks x V/16 x 64 lanes: positions
    vnm::DBuf<uint8_t> dvb;                                // per V-row block: 1 = C holds a dense sum
    std::map<int, int> shapes;                             // (b dtype << 24 | n) -> values per lane
    vnm::DBuf<uint16_t> gcn_h;                             // X W of the last GCN layer (BF16)
};

namespace vnm {

constexpr int ITEM = 64;      // row entries per work item
constexpr int GAP_MAX = 8;    // runs of rows without entries up to this long are written by the
                              // item walk; longer runs by the fix-up pass (one wave per row)

// ------------------------------------------------------------ value codecs

// OCP FP8 E4M3 (bias 7, no infinities, 0x7F = NaN, max 448): round to
// nearest even, saturating to +-448.
__host__ __device__ inline uint8_t f32_to_e4m3(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    const uint32_t a = u & 0x7FFFFFFFu;
    if (a > 0x7F800000u) return (uint8_t)(sign | 0x7F);  // NaN
    float af;
    memcpy(&af, &a, 4);
    if (af >= 448.0f) return (uint8_t)(sign | 0x7E);      // saturate
    const int e = (int)(a >> 23) - 127;                    // unbiased exponent
    if (e < -6) {
        // subnormal range: value = m * 2^-9, m in [0, 7]
        float q = af * 512.0f;                             // exact scaling
        // round half to even
        float r = q - floorf(q);
        uint32_t m = (uint32_t)floorf(q);
        if (r > 0.5f || (r == 0.5f && (m & 1u))) m += 1;
        return (uint8_t)(sign | m);                        // m == 8 -> smallest normal 0x08
    }
    // normal: keep 3 mantissa bits, round half to even on the 20 dropped bits
    uint32_t mant = a & 0x7FFFFFu;
    uint32_t keep = mant >> 20, rest = mant & 0xFFFFFu;
    uint32_t exp8 = (uint32_t)(e + 7);
    if (rest > 0x80000u || (rest == 0x80000u && (keep & 1u))) {
        keep += 1;
        if (keep == 8) { keep = 0; exp8 += 1; }
    }
    uint32_t code = (exp8 << 3) | keep;
    if (code > 0x7E) code = 0x7E;
    return (uint8_t)(sign | code);
}

__host__ __device__ inline float e4m3_to_f32(uint8_t c) {
    const uint32_t s = c & 0x80u, e = (c >> 3) & 0xFu, m = c & 7u;
    float v;
    if (e == 0) v = (float)m * (1.0f / 512.0f);
    else if (e == 15 && m == 7) v = __builtin_nanf("");
    else {
        uint32_t u = ((e + 120u) << 23) | (m << 20);
        memcpy(&v, &u, 4);
    }
    return s ? -v : v;
}

// ------------------------------------------------------------ SpMM device code

template <int BT>
constexpr int bsize() { return BT == VNM_DT_F32 ? 4 : BT == VNM_DT_BF16 ? 2 : 1; }

typedef float f32x2_t __attribute__((ext_vector_type(2)));

// VEC consecutive B values (VEC * sizeof(value) bytes, 2..16) of one row.
template <int BT, int VEC>
struct BRaw {
    static constexpr int BYTES = VEC * bsize<BT>();
    static constexpr int NW = BYTES < 4 ? 1 : BYTES / 4;
    uint32_t w[NW];
};

template <int BT, int VEC>
__device__ __forceinline__ BRaw<BT, VEC> zero_raw() {
    BRaw<BT, VEC> r;
#pragma unroll
    for (int i = 0; i < BRaw<BT, VEC>::NW; ++i) r.w[i] = 0;
    return r;
}

// acc += w * (decoded raw)
template <int BT, int VEC>
__device__ __forceinline__ void fma_raw(float (&acc)[VEC], float w, const BRaw<BT, VEC> &raw) {
    if constexpr (BT == VNM_DT_F32) {
#pragma unroll
        for (int i = 0; i < VEC; ++i) acc[i] = fmaf(w, __uint_as_float(raw.w[i]), acc[i]);
    } else if constexpr (BT == VNM_DT_BF16) {
#pragma unroll
        for (int i = 0; i < VEC; ++i) {
            const uint32_t x = raw.w[i / 2];
            acc[i] = fmaf(w, __uint_as_float((i & 1) ? (x & 0xFFFF0000u) : (x << 16)), acc[i]);
        }
    } else {
        // OCP e4m3 through the gfx950 conversion, two values per call
#pragma unroll
        for (int i = 0; i < VEC; i += 4) {
            const uint32_t x = raw.w[i / 4];
            const f32x2_t lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)x, false);
            acc[i] = fmaf(w, lo.x, acc[i]);
            acc[i + 1] = fmaf(w, lo.y, acc[i + 1]);
            if constexpr (VEC >= 4) {
                const f32x2_t hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)x, true);
                acc[i + 2] = fmaf(w, hi.x, acc[i + 2]);
                acc[i + 3] = fmaf(w, hi.y, acc[i + 3]);
            }
        }
    }
}

// Store a finished output row slice (plus the bias slice, when given).
template <int NV>
__device__ __forceinline__ void store_vec(float *__restrict__ cp, const float (&acc)[NV],
                                          const float *__restrict__ bias) {
    float v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = bias ? acc[i] + bias[i] : acc[i];
    if constexpr (NV % 4 == 0) {
#pragma unroll
        for (int i = 0; i < NV; i += 4) *(float4 *)(cp + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
    } else if constexpr (NV == 2) {
        *(float2 *)cp = make_float2(v[0], v[1]);
    } else {
        cp[0] = v[0];
    }
}

template <int NV>
__device__ __forceinline__ void load_vec(const float *__restrict__ cp, float (&v)[NV]) {
    if constexpr (NV % 4 == 0) {
#pragma unroll
        for (int i = 0; i < NV; i += 4) {
            const float4 x = *(const float4 *)(cp + i);
            v[i] = x.x; v[i + 1] = x.y; v[i + 2] = x.z; v[i + 3] = x.w;
        }
    } else if constexpr (NV == 2) {
        const float2 x = *(const float2 *)cp;
        v[0] = x.x; v[1] = x.y;
    } else {
        v[0] = cp[0];
    }
}

// One packed matrix as the SpMM kernels see it.
struct SpmmArgs {
    const int32_t *item_beg, *item_slot, *ent_row, *ent_col;
    const uint8_t *ent_idx, *ent_val;
    const int32_t *ent_blk;
    const float *blk_scale;
    const uint8_t *B;
    const float *b_scales;
    float *C;
    float *part;        // nslots x pn partial sums of cut rows
    const float *bias;  // nullptr or n floats added to every output row
    const uint8_t *dvb; // nullptr, or per V-row block 1 when C already holds the block's dense sums
    int items, M, ldb, ldc, pn;
    int b_rows;         // rows of B (columns of A); b_rows * ldb * sizeof(value) < 2^32
    int c_rows;         // rows of C (of A)
    int V;
};

// VEC values of one B row for this lane, by a buffer load whose row offset
// (soff) is a scalar and whose column offset (voff) is fixed per lane.
template <int BT, int VEC>
__device__ __forceinline__ BRaw<BT, VEC> load_row(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    BRaw<BT, VEC> r;
    constexpr int BY = BRaw<BT, VEC>::BYTES;
    if constexpr (BY == 16) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)soff, 0);
        r.w[0] = t[0]; r.w[1] = t[1]; r.w[2] = t[2]; r.w[3] = t[3];
    } else if constexpr (BY == 8) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, 0);
        r.w[0] = t[0]; r.w[1] = t[1];
    } else if constexpr (BY == 4) {
        r.w[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0);
    } else {
        r.w[0] = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)voff, (int)soff, 0);
    }
    return r;
}

// Value i of the value array as a float (per lane).
template <int DT>
__device__ __forceinline__ float load_aval(const uint8_t *__restrict__ p, int i) {
    if constexpr (DT == VNM_DT_F32) return ((const float *)p)[i];
    else if constexpr (DT == VNM_DT_BF16) return bf16_to_f32(((const uint16_t *)p)[i]);
    else if constexpr (DT == VNM_DT_F16) return __half2float(((const __half *)p)[i]);
    else return e4m3_to_f32(p[i]);
}

__device__ __forceinline__ int lane_bcast(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ __forceinline__ float lane_bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// C = A_vnm B for one work item (`cnt` <= ITEM row entries starting at
// `beg`) and column tile `tile` (64 * VEC columns), one wave.  Lane k first
// resolves entry k: its row, its scaled values and the byte offsets of its
// B rows (column * ldb * sizeof(B value); -1 for an absent second value).
// The wave then walks the entries, U at a time, with each entry's data
// broadcast by readlane: every lane gathers VEC values of the entry's B rows
// by buffer loads whose row offset is that broadcast scalar, and owns VEC
// columns of the output row.  A row is stored when the row changes, and the
// rows that keep nothing between it and the next (up to the next item's first
// row `next_row`; item 0 also from row 0) get the bias or zero.  Items
// hold whole rows, except chunks of rows longer than ITEM entries: such a
// chunk is the first row of its item, and its sum goes to the item's
// partial-sum slot (slot >= 0) instead of C; the fix-up pass adds a cut
// row's slots in item order.  So every output value is one fixed sequence of
// FP32 operations, whatever the launch shape.  PIPE: two register sets of U
// entries, the gathers of the next group in flight while the current one is
// folded in (the persistent batched kernel runs faster this way; the
// one-matrix kernel does not).
template <int AT, int BT, int VEC, bool PIPE = false>
__device__ __forceinline__ void spmm_item(const int32_t *__restrict__ erow, const int32_t *__restrict__ ecol,
                                          const uint8_t *__restrict__ eidx, const uint8_t *__restrict__ evals,
                                          const int32_t *__restrict__ eblk, const float *__restrict__ blk_scale,
                                          const float *__restrict__ bsc, const uint8_t *__restrict__ B,
                                          float *__restrict__ C, float *__restrict__ part,
                                          const float *__restrict__ bias, const uint8_t *__restrict__ dvb,
                                          int V, int beg, int cnt, int slot, int lead_row, int next_row, int M,
                                          int ldb, int ldc, int pn, int b_rows, int tile) {
    constexpr int SB = bsize<BT>();
    // entries per group: up to 8; pipelined, 16 VGPRs of gathered values
    // per register set
    constexpr int NW = BRaw<BT, VEC>::NW;
    constexpr int U = PIPE ? (NW >= 4 ? 2 : 8 / NW) : (BRaw<BT, VEC>::BYTES >= 16 ? 4 : 8);
    const int lane = threadIdx.x & 63;
    const int rowb = ldb * SB;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)B, 0, (int)((uint32_t)b_rows * (uint32_t)rowb), 0x00020000);
    int rr = 0, off0 = 0, off1 = -1;
    float a0 = 0.f, a1 = 0.f;
    if (lane < cnt) {
        const int e = beg + lane;
        rr = erow[e];
        const int c0 = ecol[e];
        const uint32_t code = eidx[e];
        float sc = 1.f;
        if constexpr (AT == VNM_DT_FP8) sc = blk_scale[eblk[e]];
        a0 = load_aval<AT>(evals, 2 * e) * sc;
        if constexpr (BT == VNM_DT_FP8) a0 *= bsc[c0 >> 7];
        off0 = c0 * rowb;
        if (code & 0x10u) {
            const int c1 = (c0 & ~(M - 1)) | (int)(code & 15u);
            a1 = load_aval<AT>(evals, 2 * e + 1) * sc;
            if constexpr (BT == VNM_DT_FP8) a1 *= bsc[c1 >> 7];
            off1 = c1 * rowb;
        }
    }
    const int first_row = lane_bcast(rr, 0);
    // entries that start a new row (entry 0 does not: cur starts there)
    const int prev_rr = __shfl_up(rr, 1, 64);
    const uint64_t starts = __ballot(lane > 0 && lane < cnt && rr != prev_rr);
    const int col0 = tile * (64 * VEC) + lane * VEC;
    const int voff = col0 * SB;
    const float *bias_c = bias ? bias + col0 : nullptr;
    float acc[VEC];
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc[i] = 0.f;
    int cur = first_row;
    // rows whose V-row block has a dense part hold its sum already (written by
    // the dense kernel): row-entry sums add to it
    auto dense_row = [&](int r) { return dvb != nullptr && dvb[r / V] != 0; };
    // rows [r0, r1) -- a maximal run without row entries -- get the bias or
    // zero (plus a dense sum); runs longer than GAP_MAX are the fix-up
    // pass's, so no wave walks thousands of empty rows
    auto fill = [&](int r0, int r1) {
        if (r1 - r0 > GAP_MAX) return;
        float z[VEC];
#pragma unroll
        for (int i = 0; i < VEC; ++i) z[i] = 0.f;
        for (int r = r0; r < r1; ++r) {
            float *cp = C + (size_t)r * ldc + col0;
            if (dense_row(r)) {
                if (bias_c) {
                    float d[VEC];
                    load_vec<VEC>(cp, d);
                    store_vec<VEC>(cp, d, bias_c);
                }
            } else {
                store_vec<VEC>(cp, z, bias_c);
            }
        }
    };
    fill(lead_row, first_row);
    // the finished row `cur`: to its slot when it is the item's cut first row
    auto flush = [&]() {
        if (slot >= 0 && cur == first_row) {
            store_vec<VEC>(part + (size_t)slot * pn + col0, acc, nullptr);
        } else {
            float *cp = C + (size_t)cur * ldc + col0;
            if (dense_row(cur)) {
                float d[VEC];
                load_vec<VEC>(cp, d);
#pragma unroll
                for (int i = 0; i < VEC; ++i) acc[i] += d[i];
            }
            store_vec<VEC>(cp, acc, bias_c);
        }
    };
    // U entries: broadcast, gather, then fold in (flushing on a row change);
    // CHECK only for the item's last, partial group
    auto group = [&](int k, auto check) {
        constexpr bool CHECK = decltype(check)::value;
        BRaw<BT, VEC> raw[U][2];
        float w[U][2];
        int o1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k + u;  // wave-uniform
            const bool ok = !CHECK || kk < cnt;
            const int ks = kk & 63;
            w[u][0] = ok ? lane_bcast(a0, ks) : 0.f;
            raw[u][0] = ok ? load_row<BT, VEC>(rs, voff, lane_bcast(off0, ks)) : zero_raw<BT, VEC>();
            o1[u] = ok ? lane_bcast(off1, ks) : -1;
            w[u][1] = o1[u] >= 0 ? lane_bcast(a1, ks) : 0.f;
            raw[u][1] = o1[u] >= 0 ? load_row<BT, VEC>(rs, voff, o1[u]) : zero_raw<BT, VEC>();
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k + u;
            if ((starts >> (kk & 63)) & 1ull) {  // wave-uniform: a new row starts here
                flush();
#pragma unroll
                for (int i = 0; i < VEC; ++i) acc[i] = 0.f;
                const int nxt = lane_bcast(rr, kk & 63);
                fill(cur + 1, nxt);
                cur = nxt;
            }
            if (!CHECK || kk < cnt) {
                fma_raw<BT, VEC>(acc, w[u][0], raw[u][0]);
                if (o1[u] >= 0) fma_raw<BT, VEC>(acc, w[u][1], raw[u][1]);
            }
        }
    };
    if constexpr (!PIPE) {
        const int full = cnt & ~(U - 1);
        for (int k = 0; k < full; k += U) group(k, std::false_type{});
        if (full < cnt) group(full, std::true_type{});
    } else {
        // the same gathers and folds as group(), split so that a group's
        // loads are issued before the previous group is folded in; entries
        // past cnt gather and fold nothing
        using Raw = BRaw<BT, VEC>[U][2];
        auto issue = [&](int k, Raw &raw, auto check) {
            constexpr bool CHECK = decltype(check)::value;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k + u;  // wave-uniform
                const bool ok = !CHECK || kk < cnt;
                const int ks = kk & 63;
                raw[u][0] = ok ? load_row<BT, VEC>(rs, voff, lane_bcast(off0, ks)) : zero_raw<BT, VEC>();
                const int o1 = ok ? lane_bcast(off1, ks) : -1;
                raw[u][1] = o1 >= 0 ? load_row<BT, VEC>(rs, voff, o1) : zero_raw<BT, VEC>();
            }
        };
        auto issue_any = [&](int k, Raw &raw) {
            if (k + U <= cnt) issue(k, raw, std::false_type{});
            else issue(k, raw, std::true_type{});
        };
        auto fold = [&](int k, const Raw &raw) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k + u;
                const int ks = kk & 63;
                if ((starts >> ks) & 1ull) {  // wave-uniform: a new row starts here
                    flush();
#pragma unroll
                    for (int i = 0; i < VEC; ++i) acc[i] = 0.f;
                    const int nxt = lane_bcast(rr, ks);
                    fill(cur + 1, nxt);
                    cur = nxt;
                }
                if (kk < cnt) {
                    fma_raw<BT, VEC>(acc, lane_bcast(a0, ks), raw[u][0]);
                    if (lane_bcast(off1, ks) >= 0) fma_raw<BT, VEC>(acc, lane_bcast(a1, ks), raw[u][1]);
                }
            }
        };
        Raw ra, rb;
        issue_any(0, ra);
        for (int k = 0;; k += 2 * U) {
            if (k + U < cnt) issue_any(k + U, rb);
            fold(k, ra);
            if (k + U >= cnt) break;
            if (k + 2 * U < cnt) issue_any(k + 2 * U, ra);
            fold(k + U, rb);
            if (k + 2 * U >= cnt) break;
        }
    }
    flush();
    fill(cur + 1, next_row);
}

// Work item `item` (wave-uniform), column tile `tile` of one packed matrix.
template <int AT, int BT, int VEC, bool PIPE = false>
__device__ __forceinline__ void spmm_unit(const SpmmArgs &a, int item, int tile) {
    const int beg = a.item_beg[item];
    const int end = a.item_beg[item + 1];
    // rows that keep nothing: item 0 fills those before its first row, every
    // item those between its rows and after its last, up to the next item's
    // first row (the last item: up to the row count)
    const int lead = item == 0 ? 0 : a.c_rows;
    const int next_row = item + 1 < a.items ? a.ent_row[end] : a.c_rows;
    spmm_item<AT, BT, VEC, PIPE>(a.ent_row, a.ent_col, a.ent_idx, a.ent_val, a.ent_blk, a.blk_scale,
                                 a.b_scales, a.B, a.C, a.part, a.bias, a.dvb, a.V, beg, end - beg,
                                 a.item_slot[item], lead, next_row, a.M, a.ldb, a.ldc, a.pn, a.b_rows, tile);
}

// Fix-up of one listed row (a wave): the sum of its `ns` partial-sum slots
// from `s0` on, in slot order, plus the row's dense sum when it has one,
// plus the bias; a matrix without work items lists every row with ns = 0.  The slots' loads go out
// 16 at a time before they are added (in order), so a row of many chunks
// waits for a few round trips rather than one per slot.
__device__ __forceinline__ void fix_row(float *__restrict__ crow, const float *__restrict__ part, int pn, int s0,
                                        int ns, int n, const float *__restrict__ bias, bool dense) {
    constexpr int FB = 16;
    const int lane = threadIdx.x & 63;
    for (int c = lane * 4; c < n; c += 256) {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        const float *pc = part + (size_t)s0 * pn + c;
        for (int j0 = 0; j0 < ns; j0 += FB) {
            float4 v[FB];
#pragma unroll
            for (int j = 0; j < FB; ++j)
                v[j] = j0 + j < ns ? *(const float4 *)(pc + (size_t)(j0 + j) * pn) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int j = 0; j < FB; ++j) {
                if (j0 + j < ns) {
                    acc.x += v[j].x;
                    acc.y += v[j].y;
                    acc.z += v[j].z;
                    acc.w += v[j].w;
                }
            }
        }
        if (dense) {  // the row's dense sum, written by the dense kernel
            const float4 d = *(const float4 *)(crow + c);
            acc.x += d.x;
            acc.y += d.y;
            acc.z += d.z;
            acc.w += d.w;
        }
        if (bias) {
            const float4 b = *(const float4 *)(bias + c);
            acc.x += b.x;
            acc.y += b.y;
            acc.z += b.z;
            acc.w += b.w;
        }
        *(float4 *)(crow + c) = acc;
    }
}

// ------------------------------------------------------------ host helpers (pack.hip)

int ensure_part(vnm_packed *p, int n);
size_t scratch_rows(const vnm_packed *p);
int bsize_of(int bt);
bool vec_ok(int bt, int n, int vec);
SpmmArgs packed_args(const vnm_packed *p, const void *B, int ldb, const float *bs, float *C, int ldc,
                     const float *bias, float *part, int pn);
int spmm(vnm_packed *p, const void *B, int ldb, int n, int bt, const float *bs, float *C, int ldc,
         const float *bias);
int spmm_check(const vnm_packed *p, const void *B, int ldb, int n, int bt, const float *bs, const float *C, int ldc);
int spmm_run(vnm_packed *p, int vec, const void *B, int ldb, int n, int bt, const float *bs, float *C, int ldc,
             const float *bias, float *part, bool timed);
int shape_key(int bt, int n);
int build_dense(vnm_packed *p, int dense_min, int U, int nb, int nblk, const int32_t *run_blk,
                const int32_t *nkeep, const int32_t *blk_first, const uint64_t *bkey, const uint64_t *keys,
                const int32_t *rstart, const float *kv, const uint8_t *kidx, DBuf<int32_t> &dmap);
int spmm_dense(vnm_packed *p, const uint16_t *B, int ldb, int n, float *C, int ldc, float *dslot, hipStream_t st,
               bool timed);

}  // namespace vnm
