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
    // the SpMM's value stream: every value of the row entries, row-major, as
    // (row, B row, FP32 weight), and VTAIL zero records past its end
    int32_t nval = 0;
    vnm::DBuf<int32_t> vrow, vcol;
    vnm::DBuf<float> vw;
    // SpMM work items: whole rows of at most ITEM values, and rows longer
    // than ITEM cut into ITEM-value chunks whose partial sums go to slots
    int32_t nitems = 0, nslots = 0, nfix = 0;
    vnm::DBuf<int32_t> item_beg;                           // nitems + 1 value offsets
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

constexpr int ITEM = 64;      // values per work item
constexpr int SBLK = 8;       // values per block of the SpMM's item walk
constexpr int VTAIL = SBLK;   // zero records past the end of the value stream
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

// store_vec for rows of C: nontemporal stores, so that the output stream
// does not push the gathered B rows out of L2 (the same values land)
template <int NV>
__device__ __forceinline__ void store_vec_nt(float *__restrict__ cp, const float (&acc)[NV],
                                             const float *__restrict__ bias) {
    float v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = bias ? acc[i] + bias[i] : acc[i];
    if constexpr (NV % 4 == 0) {
        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int i = 0; i < NV; i += 4)
            __builtin_nontemporal_store(f4{v[i], v[i + 1], v[i + 2], v[i + 3]}, (f4 *)(cp + i));
    } else if constexpr (NV == 2) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(f2{v[0], v[1]}, (f2 *)cp);
    } else {
        __builtin_nontemporal_store(v[0], cp);
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
    const int32_t *item_beg, *item_slot;
    const int32_t *vrow, *vcol;  // the value stream
    const float *vw;
    const uint8_t *B;
    const float *b_scales;
    float *C;
    float *part;        // nslots x pn partial sums of cut rows
    const float *bias;  // nullptr or n floats added to every output row
    const uint8_t *dvb; // nullptr, or per V-row block 1 when C already holds the block's dense sums
    int items, ldb, ldc, pn;
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

// C = A_vnm B for one work item (`cnt` <= ITEM values from `beg` on in the
// value stream) and column tile `tile` of T (64 * VEC columns), one wave.  The
// value stream holds every kept value of the row entries in row-major order
// as (row, B row, weight); the weight is the stored value times its block's
// scale (FP8 A), rounded to FP32 once at packing, as a product formed here
// would be.  Lane k reads the row of value k and a ballot marks where rows
// start; the wave then walks the item in blocks of 8 values.  A block's B
// rows and weights are scalar loads, so per value the wave spends one scalar
// multiply (the B row's byte offset), one buffer load (VEC consecutive values
// of the B row per lane, 2 to 16 bytes) and VEC FMAs that take the weight as
// a scalar operand: no per-value lane broadcasts or branches.  All 8 gathers
// of a block are in flight before its FMAs, whatever the row lengths; the
// FMAs go one row segment of the block at a time.  Each lane owns VEC
// columns of the output row and stores them when the row ends; the rows that
// keep nothing between it and the next (up to the next item's first row
// `next_row`; item 0 also from row 0) get the bias or zero.  Items hold whole
// rows, except chunks of rows longer than ITEM values: such a chunk is the
// first row of its item, and its sum goes to the item's partial-sum slot
// (slot >= 0) instead of C; the fix-up pass adds a cut row's slots in item
// order.  So every output value is one fixed sequence of FP32 operations (the
// row's values in order), whatever the launch shape.
template <int BT, int VEC>
__device__ __forceinline__ void spmm_item(const int32_t *__restrict__ vrow, const int32_t *__restrict__ vcol,
                                          const float *__restrict__ vw, const float *__restrict__ bsc,
                                          const uint8_t *__restrict__ B, float *__restrict__ C,
                                          float *__restrict__ part, const float *__restrict__ bias,
                                          const uint8_t *__restrict__ dvb, int V, int beg, int cnt, int slot,
                                          int lead_row, int next_row, int ldb, int ldc, int pn, int b_rows,
                                          int tile, int T) {
    constexpr int SB = bsize<BT>();
    const int lane = threadIdx.x & 63;
    const int rowb = ldb * SB;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)B, 0, (int)((uint32_t)b_rows * (uint32_t)rowb), 0x00020000);
    const int rr = lane < cnt ? vrow[beg + lane] : 0;
    const int first_row = lane_bcast(rr, 0);
    // values that start a new row (value 0 does not: cur starts there)
    const int prev_rr = __shfl_up(rr, 1, 64);
    const uint64_t starts = __ballot(lane > 0 && lane < cnt && rr != prev_rr);
    // the T column tiles interleave in 128-byte pieces (one L2 line of a B
    // row): tile t takes pieces t, t + T, t + 2T, ... of every row, so each
    // XCD's L2 holds its 1/T of B spread over the whole row
    constexpr int LB = VEC * SB;  // bytes per lane and B row
    auto tile_byte = [&](int b) { return ((b >> 7) * T + tile) * 128 + (b & 127); };
    // 8 bytes per lane: two B rows per load instruction, 16 bytes per lane --
    // lanes 0-31 read the tile of one row, lanes 32-63 that of the next, and
    // two lane-half swaps (v_permlane32_swap) then give every lane both rows'
    // 8 bytes for the columns it owns (lane i and i + 32 own the two halves
    // of bytes [16i, 16i + 16))
    constexpr bool PAIR = LB == 8;
    const int upper = lane >> 5;
    const int col0 = tile_byte(PAIR ? 16 * (lane & 31) + 8 * upper : lane * LB) / SB;
    const int voff = PAIR ? tile_byte(16 * (lane & 31)) : col0 * SB;
    const float *bias_c = bias ? bias + col0 : nullptr;
    float acc[VEC];
    int cur = first_row;
    // rows whose V-row block has a dense part hold its sum already (written by
    // the dense kernel): row sums add to it
    auto dense_row = [&](int r) { return dvb != nullptr && dvb[r / V] != 0; };
    // rows [r0, r1) -- a maximal run without values -- get the bias or zero
    // (plus a dense sum); runs longer than GAP_MAX are the fix-up pass's, so
    // no wave walks thousands of empty rows
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
                store_vec_nt<VEC>(cp, z, bias_c);
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
            store_vec_nt<VEC>(cp, acc, bias_c);
        }
    };
#pragma unroll
    for (int i = 0; i < VEC; ++i) acc[i] = 0.f;
    const int32_t *__restrict__ cp = vcol + beg;
    const float *__restrict__ wp = vw + beg;
    // the B row of one value (the weight picks up the B row's FP8 scale)
    auto gather = [&](int c, float &w) {
        if constexpr (BT == VNM_DT_FP8) w *= bsc[c >> 7];
        return load_row<BT, VEC>(rs, voff, (uint32_t)c * (uint32_t)rowb);
    };
    // the B rows of two values (PAIR), each lane's offset in full in the
    // per-lane operand (in range, like every B row)
    auto gather2 = [&](int c0, float &w0, int c1, float &w1, BRaw<BT, VEC> &r0, BRaw<BT, VEC> &r1) {
        if constexpr (BT == VNM_DT_FP8) {
            w0 *= bsc[c0 >> 7];
            w1 *= bsc[c1 >> 7];
        }
        const uint32_t o0 = __builtin_amdgcn_readfirstlane((uint32_t)c0 * (uint32_t)rowb);
        const uint32_t o1 = __builtin_amdgcn_readfirstlane((uint32_t)c1 * (uint32_t)rowb);
        const uint32_t vo = (uint32_t)voff + (upper ? o1 : o0);
        const auto t = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo, 0, 0);
        const auto s0 = __builtin_amdgcn_permlane32_swap(t[0], t[2], false, false);
        const auto s1 = __builtin_amdgcn_permlane32_swap(t[1], t[3], false, false);
        r0.w[0] = s0[0];
        r0.w[1] = s1[0];
        r1.w[0] = s0[1];
        r1.w[1] = s1[1];
    };
    // blocks of 8 values (the item's last may hold fewer, m): their B rows and
    // weights by scalar loads, all gathers in flight, then the FMAs in value
    // order, one row segment of the block at a time (a block without a row
    // start is one segment)
    constexpr int BLK = SBLK;
    for (int k0 = 0; k0 < cnt; k0 += BLK) {
        const int m = min(BLK, cnt - k0);
        int c[BLK];
        float w[BLK];
#pragma unroll
        for (int t = 0; t < BLK; ++t) {  // past the item: later items or the stream's zero tail
            c[t] = cp[k0 + t];
            w[t] = wp[k0 + t];
        }
        BRaw<BT, VEC> raw[BLK];
        if constexpr (PAIR) {
            if (m == BLK) {
#pragma unroll
                for (int t = 0; t < BLK; t += 2) gather2(c[t], w[t], c[t + 1], w[t + 1], raw[t], raw[t + 1]);
            } else {
                float wd = 0.f;
                BRaw<BT, VEC> rd;
#pragma unroll
                for (int t = 0; t < BLK; t += 2) {
                    if (t + 1 < m) gather2(c[t], w[t], c[t + 1], w[t + 1], raw[t], raw[t + 1]);
                    else if (t < m) gather2(c[t], w[t], c[t], wd, raw[t], rd);  // the last, unpaired
                }
            }
        } else if (m == BLK) {
#pragma unroll
            for (int t = 0; t < BLK; ++t) raw[t] = gather(c[t], w[t]);
        } else {
#pragma unroll
            for (int t = 0; t < BLK; ++t)
                if (t < m) raw[t] = gather(c[t], w[t]);
        }
        // values of this block that start a row (value 0's bit is clear)
        uint32_t rem = (uint32_t)(starts >> k0) & ((1u << BLK) - 1u);
        if (m == BLK && rem == 0) {
#pragma unroll
            for (int t = 0; t < BLK; ++t) fma_raw<BT, VEC>(acc, w[t], raw[t]);
            continue;
        }
        for (int t0 = 0;;) {
            // values [t0, t1) of the block belong to row `cur`
            const int t1 = rem ? (int)__builtin_ctz(rem) : m;
            const uint32_t seg = ((1u << t1) - 1u) & ~((1u << t0) - 1u);
#pragma unroll
            for (int t = 0; t < BLK; ++t)
                if ((seg >> t) & 1u) fma_raw<BT, VEC>(acc, w[t], raw[t]);
            if (!rem) break;
            // a new row starts at value k0 + t1
            flush();
#pragma unroll
            for (int i = 0; i < VEC; ++i) acc[i] = 0.f;
            const int nxt = lane_bcast(rr, k0 + t1);
            fill(cur + 1, nxt);
            cur = nxt;
            t0 = t1;
            rem &= rem - 1u;
        }
    }
    flush();
    fill(cur + 1, next_row);
}

// Work item `item` (wave-uniform), column tile `tile` of T of one packed matrix.
template <int BT, int VEC>
__device__ __forceinline__ void spmm_unit(const SpmmArgs &a, int item, int tile, int T) {
    const int beg = a.item_beg[item];
    const int end = a.item_beg[item + 1];
    // rows that keep nothing: item 0 fills those before its first row, every
    // item those between its rows and after its last, up to the next item's
    // first row (the last item: up to the row count)
    const int lead = item == 0 ? 0 : a.c_rows;
    const int next_row = item + 1 < a.items ? a.vrow[end] : a.c_rows;
    spmm_item<BT, VEC>(a.vrow, a.vcol, a.vw, a.b_scales, a.B, a.C, a.part, a.bias, a.dvb, a.V, beg, end - beg,
                       a.item_slot[item], lead, next_row, a.ldb, a.ldc, a.pn, a.b_rows, tile, T);
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
