// VENOM row reorderer on the GPU (replaces reordering.py, run by the
// reference through system() with temp.mtx / output.txt IPC, main.cu:116-193).
//
// Per iteration (reordering.py:212-255):
//   1. sort the COO positions by (row, col); runs of equal (row, col / M) give
//      the per-(row, segment) bitmaps of generate_segmat (:113-119; every
//      stored position counts, whatever its value);
//   2. V-block counts (count_invalid_venom_block[_strict], :125-152): each run
//      ORs its bitmap (and a "row has > 2 ones" flag) into the cell
//      (segment, row / V); the first run to visit a cell afterwards counts it;
//   3. the stable row sort by the dense hash lists (:162-169, :219): tokens
//      tok = sign(h) * (S - seg) * scale + h + scale / 2 (scale 512 for
//      M <= 8, 2^17 for M = 16, so |h| < scale / 2) make the dense lexicographic
//      order (absent segments hash to 0) a plain lexicographic order of the
//      rows' token lists with an end marker 0; rows are then sorted by bucket
//      refinement: at level k every row still tied with another is stably
//      re-sorted by (bucket start, k-th token) until no ties are left open;
//   4. new_id = sorted order, entry (r, c) -> (new_id[r], new_id[c]) (:171-177).
#include "graph.h"
#include "hip_common.h"
#include "sortlib.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace vnm {

namespace {

int bits_for(int64_t x) {
    int b = 1;
    while ((1ll << b) <= x) ++b;
    return b;
}

int grid_of(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65535));
}

}  // namespace

__global__ void k_pos_keys(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                           int64_t nnz, uint64_t *__restrict__ keys) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = ((uint64_t)(uint32_t)row[i] << 32) | (uint32_t)col[i];
}

__global__ void k_run_heads(const uint64_t *__restrict__ k, int64_t nnz, int M,
                            int32_t *__restrict__ head) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        int h = 1;
        if (i > 0) {
            uint64_t a = k[i], b = k[i - 1];
            h = ((a >> 32) != (b >> 32)) || ((uint32_t)a / M != (uint32_t)b / M);
        }
        head[i] = h;
    }
}

__global__ void k_run_fill(const uint64_t *__restrict__ k, const int32_t *__restrict__ head,
                           const int32_t *__restrict__ excl, int64_t nnz, int M,
                           uint32_t *__restrict__ bits, int32_t *__restrict__ rrow,
                           int32_t *__restrict__ rseg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = k[i];
        const uint32_t c = (uint32_t)key;
        const int u = excl[i] + head[i] - 1;
        atomicOr(&bits[u], 1u << (M - 1 - (c % M)));
        if (head[i]) {
            rrow[u] = (int32_t)(key >> 32);
            rseg[u] = (int32_t)(c / M);
        }
    }
}

__global__ void k_cells_or(const int32_t *__restrict__ rrow, const int32_t *__restrict__ rseg,
                           const uint32_t *__restrict__ bits, int U, int V, int nblk,
                           uint32_t *__restrict__ cells) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < U; i += gridDim.x * blockDim.x) {
        const uint32_t b = bits[i];
        const uint32_t flag = popc32(b) > 2 ? 0x10000u : 0u;  // strict: a row with > 2 ones (bitmaps <= 16 bits)
        atomicOr(&cells[(size_t)rseg[i] * nblk + rrow[i] / V], b | flag);
    }
}

__global__ void k_cells_count(const int32_t *__restrict__ rrow, const int32_t *__restrict__ rseg,
                              int U, int V, int nblk, uint32_t *__restrict__ cells,
                              int32_t *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int a = 0, b = 0;
    for (; i < U; i += gridDim.x * blockDim.x) {
        const uint32_t old = atomicOr(&cells[(size_t)rseg[i] * nblk + rrow[i] / V], 0x80000000u);
        if (!(old & 0x80000000u)) {  // first visitor counts the cell
            const bool inv = popc32(old & 0xFFFFu) > 4;
            a += inv;
            b += inv || (old & 0x10000u);
        }
    }
    // wave then block reduction
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicAdd(&out[0], a);
        if (b) atomicAdd(&out[1], b);
    }
}

__global__ void k_cells_clear(const int32_t *__restrict__ rrow, const int32_t *__restrict__ rseg,
                              int U, int V, int nblk, uint32_t *__restrict__ cells) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < U; i += gridDim.x * blockDim.x) cells[(size_t)rseg[i] * nblk + rrow[i] / V] = 0u;
}

// Row token lists: rcnt[row] = runs of the row; tok[u] for run u.
// hash (reordering.py:95-107): k with k ^ (k >> 1) == b (inverse Gray code,
// the prefix XOR of b), negated when b has more than two ones.  Runs are in
// row order, so a wave's runs of one row are contiguous lanes: the last of
// them adds the count (one atomic per row and wave instead of one per run,
// which serialises on a hub row's counter).
__global__ __launch_bounds__(256) void k_tokens(const int32_t *__restrict__ rrow, const int32_t *__restrict__ rseg,
                                                const uint32_t *__restrict__ bits, int U, int S, int scale,
                                                int32_t *__restrict__ tok, int32_t *__restrict__ rcnt) {
    const int lane = threadIdx.x & 63;
    for (int base = blockIdx.x * blockDim.x; base < U; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        const bool valid = i < U;
        const int row = valid ? rrow[i] : -1;
        if (valid) {
            const uint32_t b = bits[i];
            uint32_t k = b;
            k ^= k >> 1;
            k ^= k >> 2;
            k ^= k >> 4;
            k ^= k >> 8;
            const int h = popc32(b) <= 2 ? (int)k : -(int)k;
            const int span = S - rseg[i];
            // |h| < scale / 2: the hash orders rows at one span, the
            // signed span orders a present segment against an absent one
            tok[i] = (h > 0 ? span : -span) * scale + h + scale / 2;
        }
        const int prev = __shfl_up(row, 1, 64), next = __shfl_down(row, 1, 64);
        const uint64_t heads = __ballot(lane == 0 || row != prev);
        if (valid && (lane == 63 || row != next)) {
            const int start = 63 - __builtin_clzll(heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1)));
            atomicAdd(&rcnt[row], lane - start + 1);
        }
    }
}

__global__ void k_ref_init(int32_t *order, int32_t *bstart, int32_t *act, int m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        order[i] = i;
        bstart[i] = 0;
        act[i] = i;
    }
}

__global__ void k_ref_keys(const int32_t *__restrict__ act, int na, const int32_t *__restrict__ order,
                           const int32_t *__restrict__ bstart, const int32_t *__restrict__ rptr,
                           const int32_t *__restrict__ tok, int level,
                           uint64_t *__restrict__ kk, int32_t *__restrict__ vv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) {
        const int p = act[i];
        const int row = order[p];
        const int beg = rptr[row], n = rptr[row + 1] - beg;
        const int t = level < n ? tok[beg + level] : 0;  // 0 = end of the list
        kk[i] = ((uint64_t)(uint32_t)bstart[p] << 32) | (uint32_t)(t ^ (int)0x80000000);
        vv[i] = row;
    }
}

__global__ void k_ref_apply(const int32_t *__restrict__ act, int na, const uint64_t *__restrict__ kk,
                            const int32_t *__restrict__ vv, int32_t *__restrict__ order,
                            int32_t *__restrict__ headpos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) {
        order[act[i]] = vv[i];
        headpos[i] = (i == 0 || kk[i] != kk[i - 1]) ? act[i] : -1;
    }
}

__global__ void k_ref_next(const int32_t *__restrict__ act, int na, const uint64_t *__restrict__ kk,
                           const int32_t *__restrict__ bpos, int32_t *__restrict__ bstart,
                           uint8_t *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) {
        bstart[act[i]] = bpos[i];
        const bool tied = (i > 0 && kk[i] == kk[i - 1]) || (i + 1 < na && kk[i + 1] == kk[i]);
        const bool ended = (uint32_t)kk[i] == 0x80000000u;
        flags[i] = tied && !ended;
    }
}

__global__ void k_venom_relabel(const int32_t *__restrict__ row, const int32_t *__restrict__ col,
                                int64_t nnz, const int32_t *__restrict__ order,
                                int32_t *__restrict__ orow, int32_t *__restrict__ ocol) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        orow[i] = order[row[i]];
        ocol[i] = order[col[i]];
    }
}

__global__ void k_iota64(int32_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (int32_t)i;
}

__global__ void k_unpack_sorted(const uint64_t *__restrict__ k, const int32_t *__restrict__ idx,
                                const float *__restrict__ vin, int64_t nnz,
                                int32_t *__restrict__ row, int32_t *__restrict__ col,
                                float *__restrict__ vout, int32_t *__restrict__ dup) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        row[i] = (int32_t)(k[i] >> 32);
        col[i] = (int32_t)(uint32_t)k[i];
        vout[i] = vin[idx[i]];
        if (i > 0 && k[i] == k[i - 1]) atomicAdd(dup, 1);
    }
}

// total[i] = p[total[i]] over the ids at creation (p holds R ids).
__global__ void k_compose_v(int32_t *__restrict__ total, int64_t n, const int32_t *__restrict__ p, int32_t R) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int32_t t = total[i];
        if (t < R) total[i] = p[t];
    }
}

// ------------------------------------------------------------ host side

struct VenomWork {
    vnm_graph *g;
    int M, V, m, n, S, nblk;
    int64_t nnz;
    DBuf<uint64_t> keys, keys2;
    DBuf<int32_t> head, excl, rrow, rseg, tok, rcnt, rptr;
    DBuf<uint32_t> bits, cells;
    DBuf<int32_t> order, bstart, act, act2, vv, vv2, headpos, bpos, cnt2, row2, col2;
    DBuf<uint64_t> kk, kk2;
    DBuf<uint8_t> flags;
    int U = 0;

    int init(vnm_graph *graph, int V_) {
        g = graph;
        if (graph->col_overflow)
            return set_error(VNM_E_ARG, "entries lie past the padded column count after a relabel of a matrix "
                                        "with fewer columns than rows (the reference's segmat raises here)");
        M = g->M;
        V = V_;
        m = ((g->nrows + M - 1) / M) * M;
        n = ((g->ncols + M - 1) / M) * M;
        S = n / M;
        nblk = (m + V - 1) / V;
        nnz = g->nnz;
        const size_t e = std::max<int64_t>(nnz, 1);
        int rc;
        if ((rc = keys.alloc(e)) || (rc = keys2.alloc(e)) || (rc = head.alloc(e)) ||
            (rc = excl.alloc(e)) || (rc = rrow.alloc(e)) || (rc = rseg.alloc(e)) ||
            (rc = tok.alloc(e)) || (rc = bits.alloc(e)) || (rc = rcnt.alloc(m + 1)) ||
            (rc = rptr.alloc(m + 1)) || (rc = cells.alloc((size_t)S * nblk)) ||
            (rc = order.alloc(m)) || (rc = bstart.alloc(m)) || (rc = act.alloc(m)) ||
            (rc = act2.alloc(m)) || (rc = vv.alloc(m)) || (rc = vv2.alloc(m)) ||
            (rc = headpos.alloc(m)) || (rc = bpos.alloc(m)) || (rc = cnt2.alloc(4)) ||
            (rc = kk.alloc(m)) || (rc = kk2.alloc(m)) || (rc = flags.alloc(m)) ||
            (rc = row2.alloc(e)) || (rc = col2.alloc(e)))
            return rc;
        VNM_HIP_TRY(hipMemsetAsync(cells.p, 0, (size_t)S * nblk * 4, g->ctx->stream));
        return VNM_OK;
    }

    // Steps 1-2 on (row, col): runs, then the invalid / strict V-block counts.
    int count(const int32_t *row, const int32_t *col, int *inv, int *strict) {
        hipStream_t st = g->ctx->stream;
        VNM_HIP_TRY(hipMemsetAsync(cnt2.p, 0, 16, st));
        if (nnz == 0) {
            U = 0;
            *inv = *strict = 0;
            return VNM_OK;
        }
        const int gb = grid_of(nnz, 256);
        hipLaunchKernelGGL(k_pos_keys, dim3(gb), dim3(256), 0, st, row, col, nnz, keys.p);
        int rc = sort_keys_u64(keys.p, keys2.p, nnz, 32 + bits_for(m), st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_run_heads, dim3(gb), dim3(256), 0, st, keys2.p, nnz, M, head.p);
        if ((rc = scan_excl_i32(head.p, excl.p, nnz, st))) return rc;
        int32_t last[2];
        VNM_HIP_TRY(hipMemcpyAsync(&last[0], excl.p + nnz - 1, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(&last[1], head.p + nnz - 1, 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemsetAsync(bits.p, 0, (size_t)nnz * 4, st));
        hipLaunchKernelGGL(k_run_fill, dim3(gb), dim3(256), 0, st, keys2.p, head.p, excl.p, nnz, M,
                           bits.p, rrow.p, rseg.p);
        VNM_HIP_TRY(hipStreamSynchronize(st));
        U = last[0] + last[1];
        const int gu = grid_of(U, 256);
        hipLaunchKernelGGL(k_cells_or, dim3(gu), dim3(256), 0, st, rrow.p, rseg.p, bits.p, U, V,
                           nblk, cells.p);
        hipLaunchKernelGGL(k_cells_count, dim3(gu), dim3(256), 0, st, rrow.p, rseg.p, U, V, nblk,
                           cells.p, cnt2.p);
        hipLaunchKernelGGL(k_cells_clear, dim3(gu), dim3(256), 0, st, rrow.p, rseg.p, U, V, nblk,
                           cells.p);
        g->ctx->launches += 7;
        VNM_LAUNCH_CHECK("venom count");
        int32_t h[2];
        VNM_HIP_TRY(hipMemcpyAsync(h, cnt2.p, 8, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        *inv = h[0];
        *strict = h[1];
        return VNM_OK;
    }

    // Steps 3-4: order[] = rows in stable hash-list order (uses the runs of
    // the last count()).
    int sort_rows() {
        hipStream_t st = g->ctx->stream;
        int rc;
        VNM_HIP_TRY(hipMemsetAsync(rcnt.p, 0, (size_t)(m + 1) * 4, st));
        if (U > 0)
            hipLaunchKernelGGL(k_tokens, dim3(grid_of(U, 256)), dim3(256), 0, st, rrow.p, rseg.p,
                               bits.p, U, S, M <= 8 ? 512 : (1 << 17), tok.p, rcnt.p);
        if ((rc = scan_excl_i32(rcnt.p, rptr.p, m + 1, st))) return rc;
        hipLaunchKernelGGL(k_ref_init, dim3((m + 255) / 256), dim3(256), 0, st, order.p, bstart.p,
                           act.p, m);
        g->ctx->launches += 2;
        int na = m;
        const int end_bit = 32 + bits_for(m);
        for (int level = 0; na > 1; ++level) {
            const int gb = (na + 255) / 256;
            hipLaunchKernelGGL(k_ref_keys, dim3(gb), dim3(256), 0, st, act.p, na, order.p,
                               bstart.p, rptr.p, tok.p, level, kk.p, vv.p);
            if ((rc = sort_pairs_u64_i32(kk.p, kk2.p, vv.p, vv2.p, na, end_bit, st))) return rc;
            hipLaunchKernelGGL(k_ref_apply, dim3(gb), dim3(256), 0, st, act.p, na, kk2.p, vv2.p,
                               order.p, headpos.p);
            if ((rc = scan_incl_max_i32(headpos.p, bpos.p, na, st))) return rc;
            hipLaunchKernelGGL(k_ref_next, dim3(gb), dim3(256), 0, st, act.p, na, kk2.p, bpos.p,
                               bstart.p, flags.p);
            if ((rc = select_flagged_i32(act.p, flags.p, act2.p, cnt2.p + 2, na, st))) return rc;
            g->ctx->launches += 6;
            VNM_LAUNCH_CHECK("venom sort");
            int32_t nn;
            VNM_HIP_TRY(hipMemcpyAsync(&nn, cnt2.p + 2, 4, hipMemcpyDeviceToHost, st));
            VNM_HIP_TRY(hipStreamSynchronize(st));
            na = nn;
            std::swap(act.p, act2.p);
        }
        return VNM_OK;
    }
};

int venom_eval(vnm_graph *g, int V, int32_t *inv, int32_t *inv_strict) {
    VenomWork w;
    int rc = w.init(g, V);
    if (rc) return rc;
    int a, b;
    if ((rc = w.count(g->row.p, g->col.p, &a, &b))) return rc;
    *inv = a;
    *inv_strict = b;
    return VNM_OK;
}

// Exact value of a float after "%.2f" -> strtod -> "%.1e" -> strtof.  The
// "%.2f" step is done exactly (v * 100 is exact in double; printf rounds the
// exact value half-to-even, as nearbyint does); the "%.1e" step goes through
// snprintf once per distinct intermediate value.
static void quantize_values(std::vector<float> &vals) {
    std::unordered_map<uint64_t, float> memo;
    for (float &v : vals) {
        double d;
        if (std::fabs(v) >= 16777216.0f || !std::isfinite(v)) {
            d = (double)v;  // integral: "%.2f" prints it exactly
        } else {
            const double q = std::nearbyint((double)v * 100.0);
            d = q / 100.0;
        }
        uint64_t key;
        memcpy(&key, &d, 8);
        auto it = memo.find(key);
        if (it != memo.end()) {
            v = it->second;
            continue;
        }
        char buf[64];
        snprintf(buf, sizeof buf, "%.1e", d);
        float r = strtof(buf, nullptr);
        memo.emplace(key, r);
        v = r;
    }
}

// The C++ reload of temp.mtx after the Python step (main.cu:268): values
// through the text round trip, then the loader's (row, col) order.
int reload_sort(vnm_graph *g, bool quantize) {
    hipStream_t st = g->ctx->stream;
    const int64_t nnz = g->nnz;
    g->epoch++;
    // mmwrite(shape=(n, n)) with n the padded column count (reordering.py:180)
    const int M = g->M;
    const int n = ((g->ncols + M - 1) / M) * M;
    if (nnz == 0) {
        g->nrows = n;
        g->ncols = n;
        return VNM_OK;
    }
    int rc;
    DBuf<uint64_t> k1, k2;
    DBuf<int32_t> i1, i2, dup;
    DBuf<float> v2;
    if ((rc = k1.alloc(nnz)) || (rc = k2.alloc(nnz)) || (rc = i1.alloc(nnz)) ||
        (rc = i2.alloc(nnz)) || (rc = v2.alloc(nnz)) || (rc = dup.alloc(1)))
        return rc;
    const int gb = grid_of(nnz, 256);
    hipLaunchKernelGGL(k_pos_keys, dim3(gb), dim3(256), 0, st, g->row.p, g->col.p, nnz, k1.p);
    hipLaunchKernelGGL(k_iota64, dim3(gb), dim3(256), 0, st, i1.p, nnz);
    if ((rc = sort_pairs_u64_i32(k1.p, k2.p, i1.p, i2.p, nnz, 64, st))) return rc;
    if (g->col_overflow)
        return set_error(VNM_E_ARG, "entries lie past the padded column count; the reference's "
                                    "coo_matrix(shape=(n, n)) raises here");
    if (g->nrows > n) {
        // coo_matrix(shape=(n, n)) refuses a row index >= n; the last sorted
        // key holds the largest row.  The matrix is left as it is.
        uint64_t last = 0;
        VNM_HIP_TRY(hipMemcpyAsync(&last, k2.p + nnz - 1, 8, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        if ((int64_t)(last >> 32) >= n)
            return set_error(VNM_E_ARG, "row %lld exceeds the %d x %d shape the matrix is written with "
                                        "(the reference's coo_matrix raises here)", (long long)(last >> 32), n, n);
    }
    g->nrows = n;
    g->ncols = n;
    if (quantize) {
        std::vector<float> v(nnz);
        VNM_HIP_TRY(hipMemcpyAsync(v.data(), g->val.p, nnz * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        quantize_values(v);
        VNM_HIP_TRY(hipMemcpyAsync(g->val.p, v.data(), nnz * 4, hipMemcpyHostToDevice, st));
    }
    VNM_HIP_TRY(hipMemsetAsync(dup.p, 0, 4, st));
    hipLaunchKernelGGL(k_unpack_sorted, dim3(gb), dim3(256), 0, st, k2.p, i2.p, g->val.p, nnz,
                       g->row.p, g->col.p, v2.p, dup.p);
    g->ctx->launches += 3;
    VNM_LAUNCH_CHECK("reload sort");
    std::swap(g->val.p, v2.p);
    int32_t ndup = 0;
    VNM_HIP_TRY(hipMemcpyAsync(&ndup, dup.p, 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    if (ndup > 0) {
        // repeated positions only follow a relabel that was not one-to-one:
        // apply the loader's exact duplicate rule on the host
        HostCoo c;
        c.nrows = g->nrows;
        c.ncols = g->ncols;
        c.row.resize(nnz); c.col.resize(nnz); c.val.resize(nnz);
        VNM_HIP_TRY(hipMemcpyAsync(c.row.data(), g->row.p, nnz * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(c.col.data(), g->col.p, nnz * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(c.val.data(), g->val.p, nnz * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        Tuples t;
        t.nrows = c.nrows; t.ncols = c.ncols; t.symmetric = false;
        t.row = c.row; t.col = c.col; t.val = c.val;
        HostCoo o;
        loader_order(t, 0, o, 0);
        g->nnz = (int64_t)o.row.size();
        VNM_HIP_TRY(hipMemcpyAsync(g->row.p, o.row.data(), g->nnz * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemcpyAsync(g->col.p, o.col.data(), g->nnz * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemcpyAsync(g->val.p, o.val.data(), g->nnz * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
    }
    return VNM_OK;
}

static int check_venom_args(vnm_graph *g, int V) {
    if (V <= 0) return set_error(VNM_E_ARG, "V must be positive (got %d)", V);
    if (g->M != 4 && g->M != 8 && g->M != 16)
        return set_error(VNM_E_ARG, "VENOM on the GPU supports M = 4, 8 or 16");
    // row tokens are (span * scale + hash) in 32 bits: scale 2^17 for M = 16
    if (g->M == 16 && (g->ncols + 15) / 16 >= (1 << 14) - 1)
        return set_error(VNM_E_ARG, "VENOM with M = 16 supports at most %d columns", ((1 << 14) - 2) * 16);
    return VNM_OK;
}

int venom_reorder(vnm_graph *g, int V, int maxiter, vnm_venom_stats *st) {
    g->epoch++;
    VenomWork w;
    int rc = w.init(g, V);
    if (rc) return rc;
    if (w.n > w.m)
        return set_error(VNM_E_ARG, "VENOM reorder needs ncols <= nrows (new_id is indexed by column)");
    hipStream_t s = g->ctx->stream;
    int init_inv, init_strict;
    if ((rc = w.count(g->row.p, g->col.p, &init_inv, &init_strict))) return rc;
    int inv = init_inv, prev = init_inv, prev_strict = init_strict, round = 1;
    const int MAXITER = init_inv > maxiter ? init_inv : maxiter;  // reordering.py:209-210
    int32_t *cur_r = g->row.p, *cur_c = g->col.p, *nxt_r = w.row2.p, *nxt_c = w.col2.p;
    int final_round = -1;
    while (inv > 0 && round <= MAXITER) {
        // one iteration of reordering.py's loop: row keys, stable sort,
        // relabel, recount (the host reads the counts)
        KernelTimer kt(g->ctx, "venom_iteration");
        if ((rc = w.sort_rows())) return rc;
        hipLaunchKernelGGL(k_venom_relabel, dim3(grid_of(g->nnz, 256)), dim3(256), 0, s, cur_r,
                           cur_c, g->nnz, w.order.p, nxt_r, nxt_c);
        g->ctx->launches++;
        if (w.n < w.m) {
            // new_id renames columns too: with more padded rows than columns
            // one can land past the n // M segments, where generate_segmat
            // raises IndexError (reordering.py:113-118).  The matrix keeps
            // the last kept round.
            int32_t mr, mc;
            if ((rc = max_ids(g, nxt_r, nxt_c, &mr, &mc))) return rc;
            if (mc >= w.n) {
                if (cur_r != g->row.p) {
                    VNM_HIP_TRY(hipMemcpyAsync(g->row.p, cur_r, g->nnz * 4, hipMemcpyDeviceToDevice, s));
                    VNM_HIP_TRY(hipMemcpyAsync(g->col.p, cur_c, g->nnz * 4, hipMemcpyDeviceToDevice, s));
                    VNM_HIP_TRY(hipStreamSynchronize(s));
                }
                return set_error(VNM_E_ARG, "VENOM round %d moves column %d past the %d padded columns "
                                            "(the reference's segmat raises here)", round, mc, w.n);
            }
        }
        int strict;
        rc = w.count(nxt_r, nxt_c, &inv, &strict);
        kt.stop();
        if (rc) return rc;
        if (inv > prev) {  // reordering.py:242-246: keep the previous positions
            final_round = round;
            break;
        }
        hipLaunchKernelGGL(k_compose_v, dim3((unsigned)((g->perm_total.n + 255) / 256)), dim3(256), 0, s,
                           g->perm_total.p, (int64_t)g->perm_total.n, w.order.p, w.m);
        g->ctx->launches++;
        prev = inv;
        prev_strict = strict;
        std::swap(cur_r, nxt_r);
        std::swap(cur_c, nxt_c);
        ++round;
    }
    if (cur_r != g->row.p) {
        VNM_HIP_TRY(hipMemcpyAsync(g->row.p, cur_r, g->nnz * 4, hipMemcpyDeviceToDevice, s));
        VNM_HIP_TRY(hipMemcpyAsync(g->col.p, cur_c, g->nnz * 4, hipMemcpyDeviceToDevice, s));
    }
    VNM_HIP_TRY(hipStreamSynchronize(s));
    st->init_inv = init_inv;
    st->init_inv_strict = init_strict;
    st->final_inv = prev;
    st->final_inv_strict = prev_strict;
    st->rounds = final_round >= 0 ? final_round : round - 1;  // :245 vs :265
    return VNM_OK;
}

}  // namespace vnm

using namespace vnm;

extern "C" {

int vnm_venom_eval(vnm_graph *g, int V, int32_t *inv, int32_t *inv_strict) {
    if (!g || !inv || !inv_strict) return set_error(VNM_E_ARG, "null argument");
    int rc = check_venom_args(g, V);
    if (rc) return rc;
    return venom_eval(g, V, inv, inv_strict);
}

int vnm_venom_reorder(vnm_graph *g, int V, int maxiter, vnm_venom_stats *st) {
    if (!g || !st) return set_error(VNM_E_ARG, "null argument");
    int rc = check_venom_args(g, V);
    if (rc) return rc;
    if ((rc = venom_reorder(g, V, maxiter, st))) return rc;
    return reload_sort(g, true);
}

}  // extern "C"
