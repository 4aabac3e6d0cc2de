// 2:4 sparse-MFMA (SMFMAC) part of the V:N:M SpMM, for dense V x M blocks.
//
// Replaces the 2:4 cuSPARSELt matmul of the reference (kernel/spmm.cu:
// 402-469, cusparseLtMatmul over the pruned, compressed A) for the part of
// A where the hardware's 2:4 sparse matrix core pays: V x M blocks that keep
// many values.  A stored block keeps 4 of its M columns and every row at most
// 2 of those 4 -- exactly the 2:4 structure v_smfmac_f32_16x16x64_bf16 takes
// for one group of 4 K positions.  So 16 blocks of one V-row block, side by
// side, form the K = 64 of one SMFMAC, and a 16-row slice of the V rows its
// M = 16:
//
//   chunk      16 dense blocks of one V-row block (segment order; the last
//              chunk of a V-row block padded with empty blocks); K position
//              4 b + j is selected column j of the chunk's block b
//              (dcol: the B row, -1 = none);
//   fragment   per (chunk, slice of 16 rows), the SMFMAC A operand in its
//              lane layout: lane l holds row l % 16 of the slice and the 4
//              blocks 4 (l / 16) .. + 3: per block 2 BF16 values (dfrag,
//              16 bytes per lane) and 2 x 2-bit positions among the block's
//              4 columns (didx, bits 4 g .. 4 g + 3 for block g of the lane);
//   B operand  lane l holds column l % 16 of the 16-column tile and K rows
//              8 (l / 16) .. + 7 and 32 + 8 (l / 16) .. + 7; the output
//              lane l rows 4 (l / 16) .. + 3 of the slice, column l % 16.
//
// k_spmm_dense: one workgroup per (unit of at most 2 chunks of a V-row block,
// group of NT 16-column tiles), one wave per 16-row slice.  Per chunk the workgroup
// stages the chunk's 64 B rows (NT x 16 columns, BF16) in LDS transposed --
// each column's 64 K values together in 16-byte units, so a lane's B operand
// is two 16-byte LDS reads -- and every wave issues NT SMFMACs with its
// slice's fragment.
// The V rows' sums are written to C (rows without dense values get zero) by
// the first unit of the V-row block, the other units' sums to dense slots
// that k_dense_combine adds to C in unit order; the row-entry kernel then adds the rows' sparse-block sums, the bias
// and (for rows cut into chunks) the fix-up pass -- a fixed order of FP32
// operations, so the product stays bit-identical run to run.
//
// A block is dense when it keeps at least dense_min values (vnm_pack_ex):
// it then costs its 4 B rows for the whole V-row block instead of one B row
// per kept value.  BF16 matrix values and BF16 B; V a multiple of 16.
#include "graph.h"
#include "hip_common.h"
#include "packed.h"

#include <algorithm>
#include <vector>

namespace vnm {

constexpr int UNIT_CHUNKS = 2;  // chunks per dense work unit

typedef short bf16x8_t __attribute__((ext_vector_type(8)));
typedef short bf16x16_t __attribute__((ext_vector_type(16)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ packing

__global__ void k_blk_kept(const int32_t *__restrict__ run_blk, const int32_t *__restrict__ nkeep, int U,
                           int32_t *__restrict__ blk_kept) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < U && nkeep[u] > 0) atomicAdd(&blk_kept[run_blk[u]], nkeep[u]);
}

__global__ void k_blk_vb(const int32_t *__restrict__ blk_first, const uint64_t *__restrict__ bkey, int nb,
                         int nblk, int32_t *__restrict__ blk_vb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) blk_vb[b] = (int32_t)((bkey[blk_first[b]] >> 32) % (uint64_t)nblk);
}

// One run (a row of a block) of a dense block: its kept values into the
// fragment of its (chunk, slice), at lane (row % 16) + 16 (slot / 4), word
// slot % 4, with ascending positions (a single value pairs with a zero at
// another position, as the 2:4 operand has exactly two per group).
__global__ void k_dense_frag(const int32_t *__restrict__ run_blk, const int32_t *__restrict__ dmap,
                             const uint64_t *__restrict__ keys, const int32_t *__restrict__ rstart,
                             const float *__restrict__ kv, const uint8_t *__restrict__ kidx,
                             const int32_t *__restrict__ nkeep, int U, int V, uint32_t *__restrict__ frag,
                             uint32_t *__restrict__ idx) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= U || nkeep[u] == 0) return;
    const int m = dmap[run_blk[u]];
    if (m < 0) return;
    const int chunk = m >> 4, slot = m & 15;
    const int row = (int)(keys[rstart[u]] >> 32);
    const int local = row % V;
    const int S = V / 16;
    const int lane = (local & 15) + 16 * (slot >> 2), g = slot & 3;
    const size_t f = ((size_t)chunk * S + (local >> 4)) * 64 + lane;
    const uint32_t code = kidx[u];
    float v0 = kv[2 * u], v1 = kv[2 * u + 1];
    int p0 = code & 3, p1 = (code >> 2) & 3;
    if (code & 0x10u) {
        if (p0 > p1) {
            const int t = p0; p0 = p1; p1 = t;
            const float tv = v0; v0 = v1; v1 = tv;
        }
    } else if (p0 < 3) {
        p1 = 3;
        v1 = 0.f;
    } else {
        p1 = 3;
        p0 = 2;
        v1 = v0;
        v0 = 0.f;
    }
    frag[f * 4 + g] = (uint32_t)f32_to_bf16(v0) | ((uint32_t)f32_to_bf16(v1) << 16);
    atomicOr(&idx[f], (uint32_t)(p0 | (p1 << 2)) << (4 * g));
}

// Dense part of a packed matrix: per-block kept counts decide which blocks
// are dense; those are grouped per V-row block into chunks of 16 (host pass
// over the block metadata), their runs fill the fragments, and dmap (per
// block: chunk * 16 + slot, or -1) tells the row-entry packing to leave them
// out.
int build_dense(vnm_packed *p, int dense_min, int U, int nb, int nblk, const int32_t *run_blk,
                const int32_t *nkeep, const int32_t *blk_first, const uint64_t *bkey, const uint64_t *keys,
                const int32_t *rstart, const float *kv, const uint8_t *kidx, DBuf<int32_t> &dmap) {
    hipStream_t st = p->ctx->stream;
    const int V = p->V, S = V / 16;
    int rc;
    if ((rc = dmap.alloc(std::max(nb, 1)))) return rc;
    std::vector<int32_t> hmap(std::max(nb, 1), -1);
    DBuf<int32_t> kept, vb;
    if ((rc = kept.alloc(std::max(nb, 1))) || (rc = vb.alloc(std::max(nb, 1)))) return rc;
    VNM_HIP_TRY(hipMemsetAsync(kept.p, 0, (size_t)std::max(nb, 1) * 4, st));
    if (U > 0) hipLaunchKernelGGL(k_blk_kept, dim3((U + 255) / 256), dim3(256), 0, st, run_blk, nkeep, U, kept.p);
    if (nb > 0) hipLaunchKernelGGL(k_blk_vb, dim3((nb + 255) / 256), dim3(256), 0, st, blk_first, bkey, nb, nblk, vb.p);
    VNM_LAUNCH_CHECK("dense block counts");
    std::vector<int32_t> hk(nb), hv(nb), hs(nb);
    std::vector<uint16_t> hsel(nb);
    if (nb > 0) {
        VNM_HIP_TRY(hipMemcpyAsync(hk.data(), kept.p, (size_t)nb * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(hv.data(), vb.p, (size_t)nb * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(hs.data(), p->blk_seg.p, (size_t)nb * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(hsel.data(), p->blk_sel.p, (size_t)nb * 2, hipMemcpyDeviceToHost, st));
    }
    VNM_HIP_TRY(hipStreamSynchronize(st));
    std::vector<int> dense;
    for (int b = 0; b < nb; ++b)
        if (hk[b] >= dense_min) dense.push_back(b);
    std::sort(dense.begin(), dense.end(), [&](int a, int b) {
        return hv[a] != hv[b] ? hv[a] < hv[b] : hs[a] < hs[b];
    });
    std::vector<int4> grp, multi;
    std::vector<int32_t> col;
    std::vector<uint8_t> vbf(nblk, 0);
    int nch = 0, nds = 0;
    for (size_t i = 0; i < dense.size();) {
        const int v = hv[dense[i]];
        size_t j = i;
        while (j < dense.size() && hv[dense[j]] == v) ++j;
        const int cnt = (int)(j - i), chunks = (cnt + 15) / 16;
        // units of at most UNIT_CHUNKS chunks: the first writes C, the others
        // partial sums to dense slots, added in unit order by the combine pass
        const int units = (chunks + UNIT_CHUNKS - 1) / UNIT_CHUNKS;
        for (int q = 0; q < units; ++q)
            grp.push_back(make_int4(v, nch + q * UNIT_CHUNKS, std::min(UNIT_CHUNKS, chunks - q * UNIT_CHUNKS),
                                    q == 0 ? -1 : nds + q - 1));
        if (units > 1) {
            multi.push_back(make_int4(v, nds, units - 1, 0));
            nds += units - 1;
        }
        vbf[v] = 1;
        col.resize((size_t)(nch + chunks) * 64, -1);
        for (int k = 0; k < cnt; ++k) {
            const int b = dense[i + k];
            const int chunk = nch + k / 16, slot = k % 16;
            hmap[b] = chunk * 16 + slot;
            for (int c = 0; c < 4; ++c)
                col[(size_t)chunk * 64 + slot * 4 + c] = hs[b] * p->M + ((hsel[b] >> (4 * c)) & 15);
        }
        nch += chunks;
        i = j;
    }
    p->dense_min = dense_min;
    p->ndense = (int)dense.size();
    p->nchunks = nch;
    p->ngroups = (int)grp.size();
    p->nmulti = (int)multi.size();
    p->ndslots = nds;
    VNM_HIP_TRY(hipMemcpyAsync(dmap.p, hmap.data(), hmap.size() * 4, hipMemcpyHostToDevice, st));
    if (nch == 0) return VNM_OK;
    const size_t nf = (size_t)nch * S * 64;
    if ((rc = p->dgrp.alloc(grp.size())) || (rc = p->dcol.alloc(col.size())) || (rc = p->dfrag.alloc(nf * 4)) ||
        (rc = p->didx.alloc(nf)) || (rc = p->dvb.alloc(nblk)) || (rc = p->dmulti.alloc(std::max<size_t>(multi.size(), 1))))
        return rc;
    VNM_HIP_TRY(hipMemcpyAsync(p->dgrp.p, grp.data(), grp.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    if (!multi.empty())
        VNM_HIP_TRY(hipMemcpyAsync(p->dmulti.p, multi.data(), multi.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    VNM_HIP_TRY(hipMemcpyAsync(p->dcol.p, col.data(), col.size() * 4, hipMemcpyHostToDevice, st));
    VNM_HIP_TRY(hipMemcpyAsync(p->dvb.p, vbf.data(), vbf.size(), hipMemcpyHostToDevice, st));
    VNM_HIP_TRY(hipMemsetAsync(p->dfrag.p, 0, nf * 16, st));
    VNM_HIP_TRY(hipMemsetAsync(p->didx.p, 0, nf * 4, st));
    hipLaunchKernelGGL(k_dense_frag, dim3((U + 255) / 256), dim3(256), 0, st, run_blk, dmap.p, keys, rstart, kv, kidx,
                       nkeep, U, V, p->dfrag.p, p->didx.p);
    VNM_LAUNCH_CHECK("k_dense_frag");
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

// ------------------------------------------------------------ SpMM

// Per V-row block with dense blocks (blockIdx.x) and group of NT 16-column
// tiles (blockIdx.y): the sum over its chunks of the 2:4 products, written
// to C for every row of the V rows.
// LDS position (BF16 units) of K row k of column c in the staged B tile:
// columns of 64 K values, in 16-byte units of 8 K values whose order is
// XOR-swizzled by the column's group of 8, so that lanes writing one K row of
// neighbouring column groups fall in different banks.
__device__ __forceinline__ int bt_pos(int c, int k) { return c * 64 + 8 * ((k >> 3) ^ ((c >> 3) & 7)) + (k & 7); }

// Per unit (blockIdx.x: up to UNIT_CHUNKS chunks of one V-row block) and
// group of NT 16-column tiles (blockIdx.y): the sum over its chunks of the
// 2:4 products, written for every row of the V rows -- to C for the V-row
// block's first unit, to the unit's dense slot (V x n floats) otherwise.
template <int NT>
__global__ __launch_bounds__(1024) void k_spmm_dense(const int4 *__restrict__ grp, const int32_t *__restrict__ dcol,
                                                     const uint4 *__restrict__ frag, const uint32_t *__restrict__ idx,
                                                     const uint16_t *__restrict__ B, int ldb, float *__restrict__ C,
                                                     int ldc, float *__restrict__ dslot, int n, int V, int R) {
    constexpr int NC = NT * 16;   // columns per workgroup
    __shared__ __align__(16) uint16_t bt[NC * 64];
    const int S = V / 16;
    const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int4 gd = grp[blockIdx.x];
    const int n0 = blockIdx.y * NC;
    const int kb = lane >> 4, j = lane & 15;
    f32x4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < gd.z; ++c) {
        const int chunk = gd.y + c;
        __syncthreads();  // the previous chunk's reads of bt are done
        // stage: 64 B rows x NC columns in 8-value pieces (consecutive threads
        // on consecutive pieces of one row), transposed into bt
        for (int q = threadIdx.x; q < 64 * (NC / 8); q += S * 64) {
            const int k = q / (NC / 8), piece = q % (NC / 8);
            const int col = dcol[(size_t)chunk * 64 + k];
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (col >= 0) v = *(const uint4 *)(B + (size_t)col * ldb + n0 + piece * 8);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int i = 0; i < 8; ++i) bt[bt_pos(piece * 8 + i, k)] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
        }
        __syncthreads();
        const size_t f = ((size_t)chunk * S + s) * 64 + lane;
        const uint4 av = frag[f];
        const int ai = (int)idx[f];
        bf16x8_t a;
        __builtin_memcpy(&a, &av, 16);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int cc = t * 16 + j;
            // K rows 8 kb .. + 7 and 32 + 8 kb .. + 7 of column cc
            const uint4 lo = *(const uint4 *)(bt + bt_pos(cc, 8 * kb));
            const uint4 hi = *(const uint4 *)(bt + bt_pos(cc, 32 + 8 * kb));
            bf16x16_t b;
            __builtin_memcpy(&b, &lo, 16);
            __builtin_memcpy((char *)&b + 16, &hi, 16);
            acc[t] = __builtin_amdgcn_smfmac_f32_16x16x64_bf16(a, b, acc[t], ai, 0, 0);
        }
    }
    const int l0 = s * 16 + 4 * kb;  // row within the V-row block
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = gd.x * V + l0 + i;
        if (r < R) {
            float *dst = gd.w < 0 ? C + (size_t)r * ldc : dslot + ((size_t)gd.w * V + l0 + i) * n;
#pragma unroll
            for (int t = 0; t < NT; ++t) dst[n0 + t * 16 + j] = acc[t][i];
        }
    }
}

// V-row blocks of several units: C += their dense slots, in unit order.
// Workgroup (blockIdx.x, blockIdx.y) takes 1,024 of the V x n values of
// multi-unit V-row block blockIdx.x; a thread's slot loads go out 16 at a
// time before they are added in order.
__global__ __launch_bounds__(256) void k_dense_combine(const int4 *__restrict__ multi, const float *__restrict__ dslot,
                                                       int n, int V, int R, float *__restrict__ C, int ldc) {
    constexpr int FB = 16;
    const int4 m = multi[blockIdx.x];
    const int rows = min(V, R - m.x * V);
    const int e = (blockIdx.y * 256 + threadIdx.x) * 4;
    if (e >= rows * n) return;
    const int l = e / n, c = e % n;
    float *cp = C + (size_t)(m.x * V + l) * ldc + c;
    float4 acc = *(const float4 *)cp;
    const float *sp = dslot + ((size_t)m.y * V + l) * n + c;
    for (int q0 = 0; q0 < m.z; q0 += FB) {
        float4 v[FB];
#pragma unroll
        for (int q = 0; q < FB; ++q)
            v[q] = q0 + q < m.z ? *(const float4 *)(sp + (size_t)(q0 + q) * V * n) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int q = 0; q < FB; ++q) {
            if (q0 + q < m.z) {
                acc.x += v[q].x;
                acc.y += v[q].y;
                acc.z += v[q].z;
                acc.w += v[q].w;
            }
        }
    }
    *(float4 *)cp = acc;
}

// Launch of the dense part for n columns (a multiple of 64); dslot holds
// ndslots x V x n floats.
int spmm_dense(vnm_packed *p, const uint16_t *B, int ldb, int n, float *C, int ldc, float *dslot, hipStream_t st,
               bool timed) {
    if (p->ngroups == 0) return VNM_OK;
    const int S = p->V / 16;
    KernelTimer kt(timed ? p->ctx : nullptr, "k_spmm_dense");
    const uint4 *fr = (const uint4 *)p->dfrag.p;
    if (n % 256 == 0)
        hipLaunchKernelGGL(k_spmm_dense<16>, dim3(p->ngroups, n / 256), dim3(S * 64), 0, st, p->dgrp.p, p->dcol.p,
                           fr, p->didx.p, B, ldb, C, ldc, dslot, n, p->V, p->nrows);
    else if (n % 128 == 0)
        hipLaunchKernelGGL(k_spmm_dense<8>, dim3(p->ngroups, n / 128), dim3(S * 64), 0, st, p->dgrp.p, p->dcol.p,
                           fr, p->didx.p, B, ldb, C, ldc, dslot, n, p->V, p->nrows);
    else
        hipLaunchKernelGGL(k_spmm_dense<4>, dim3(p->ngroups, n / 64), dim3(S * 64), 0, st, p->dgrp.p, p->dcol.p,
                           fr, p->didx.p, B, ldb, C, ldc, dslot, n, p->V, p->nrows);
    if (p->nmulti > 0)
        hipLaunchKernelGGL(k_dense_combine, dim3(p->nmulti, (p->V * n + 1023) / 1024), dim3(256), 0, st, p->dmulti.p,
                           dslot, n, p->V, p->nrows, C, ldc);
    kt.stop();
    p->ctx->launches += p->nmulti > 0 ? 2 : 1;
    VNM_LAUNCH_CHECK("k_spmm_dense");
    return VNM_OK;
}

}  // namespace vnm
