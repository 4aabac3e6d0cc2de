// Per-segment row tables shared by the N:M engine and the learned proposer.
//
// A table slot holds one byte per row (R bytes, 4 rows per 32-bit word): the
// M-bit bitmap of that row in one segment, MSB = the segment's first column
// (reorder/reorder.h:87-102).  A slot is filled by atomicOr from the
// segment's M column lists (CSC over run-start ids, current column s*M + b =
// run-start column inv[s*M + b]), read back once per entry, then cleared by
// the same entries, so slots are all-zero between uses.
//
// The entry passes are latency-bound (a binary search in LDS, then a
// dependent load of the entry's row): each thread resolves TAB_U entries
// before it issues their loads, so TAB_U chains are in flight at once.
#pragma once

#include "hip_common.h"

namespace vnm {

constexpr int TAB_U = 4;

// Column list of slot j, bit b: run-start column inv[seg*M + b].
struct ColSpan { int start, len; };

// Fill per-column spans for K segments (K*M <= 128 columns) and their prefix
// sums.  Must be called by all threads; returns the total entry count.
__device__ inline int table_columns(const int *sel, int K, int M, const int32_t *inv,
                             const int32_t *cptr, ColSpan *cols, int *pre) {
    const int nc = K * M;
    for (int t = threadIdx.x; t < nc; t += blockDim.x) {
        int j = t / M, b = t - j * M;
        int c0 = ld_l2(&inv[sel[j] * M + b]);
        int s = cptr[c0];
        cols[t].start = s;
        cols[t].len = cptr[c0 + 1] - s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // exclusive scan of the lengths, two columns per lane
        const int lane = threadIdx.x;
        const int l0 = 2 * lane < nc ? cols[2 * lane].len : 0;
        const int l1 = 2 * lane + 1 < nc ? cols[2 * lane + 1].len : 0;
        int v = l0 + l1;
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(v, d, 64);
            if (lane >= d) v += t;
        }
        const int ex = v - l0 - l1;
        if (2 * lane < nc) pre[2 * lane] = ex;
        if (2 * lane + 1 < nc) pre[2 * lane + 1] = ex + l0;
        if (lane == 63) pre[nc] = v;
    }
    __syncthreads();
    return pre[nc];
}

// Locate the column of entry e (0 <= e < total) by binary search on pre.
__device__ __forceinline__ int entry_col(const int *pre, int nc, int e) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t *slot_word(uint32_t *tab, int j, int R, int r) {
    return tab + (size_t)j * (size_t)(R >> 2) + (r >> 2);
}

__device__ __forceinline__ uint32_t slot_byte(uint32_t *tab, int j, int R, int r) {
    return (ld_l2(slot_word(tab, j, R, r)) >> ((r & 3) * 8)) & 0xFFu;
}

// Resolve up to TAB_U entries of this thread (entries e0, e0 + stride, ...):
// their column index t (-1 past total) and row.
__device__ __forceinline__ void table_entries(const ColSpan *cols, const int *pre, int nc, int total,
                                              const int32_t *crow, int e0, int stride,
                                              int (&t)[TAB_U], int (&r)[TAB_U]) {
#pragma unroll
    for (int u = 0; u < TAB_U; ++u) {
        const int e = e0 + u * stride;
        t[u] = e < total ? entry_col(pre, nc, e) : -1;
    }
#pragma unroll
    for (int u = 0; u < TAB_U; ++u) {
        const int e = e0 + u * stride;
        r[u] = t[u] >= 0 ? crow[cols[t[u]].start + (e - pre[t[u]])] : 0;
    }
}

// One pass over this workgroup's share (part of parts) of the entries of
// nc columns: set (CLEAR = false) or clear the row bytes of slots slot0 +
// t / M.  Entry e belongs to part (e / blockDim.x) % parts.
template <bool CLEAR>
__device__ inline void table_pass(uint32_t *tab, int slot0, int M, int R, const ColSpan *cols,
                                  const int *pre, int nc, int total, const int32_t *crow, int part,
                                  int parts) {
    const int stride = parts * blockDim.x;
    for (int e0 = part * blockDim.x + threadIdx.x; e0 < total; e0 += stride * TAB_U) {
        int t[TAB_U], r[TAB_U];
        table_entries(cols, pre, nc, total, crow, e0, stride, t, r);
#pragma unroll
        for (int u = 0; u < TAB_U; ++u) {
            if (t[u] < 0) continue;
            uint32_t *w = slot_word(tab, slot0 + t[u] / M, R, r[u]);
            if constexpr (CLEAR) st_l2(w, 0u);
            else atomicOr(w, (1u << (M - 1 - t[u] % M)) << ((r[u] & 3) * 8));
        }
    }
}

// ---- row hash in LDS, for segment sets of at most HASH_MAX entries
//
// The global slots above cost R bytes per workgroup and one scattered
// atomic, read and clear per entry in a table far larger than the L2; a set
// of a few thousand entries touches at most that many rows, which an
// open-addressing hash in LDS holds instead: key = row, value = the row's
// bitmaps (byte j = slot j, as in the global slots).  Probes are bounded by
// the capacity, so every loop ends.

constexpr int HASH_BITS = 12;
constexpr int HASH_CAP = 1 << HASH_BITS;  // slots (32 KiB of LDS with the values)
constexpr int HASH_MAX = HASH_CAP / 2;    // entries of a set that uses the hash
constexpr uint32_t HASH_EMPTY = 0xFFFFFFFFu;
constexpr int TILE_ROWS = HASH_CAP * 8;    // rows of an LDS byte tile in the hash's storage

struct RowHash {
    uint32_t key[HASH_CAP];
    uint32_t val[HASH_CAP];
};

__device__ __forceinline__ uint32_t hash_home(uint32_t r) { return (r * 0x9E3779B1u) >> (32 - HASH_BITS); }

// All slots empty; call with every thread, synchronise before use.
__device__ inline void hash_reset(RowHash &H) {
    for (int i = threadIdx.x; i < HASH_CAP; i += blockDim.x) {
        H.key[i] = HASH_EMPTY;
        H.val[i] = 0u;
    }
}

__device__ __forceinline__ void hash_or(RowHash &H, uint32_t r, uint32_t bits) {
    uint32_t h = hash_home(r);
    for (int p = 0; p < HASH_CAP; ++p) {
        const uint32_t k = atomicCAS(&H.key[h], HASH_EMPTY, r);
        if (k == HASH_EMPTY || k == r) {
            atomicOr(&H.val[h], bits);
            return;
        }
        h = (h + 1) & (HASH_CAP - 1);
    }
}

// Bitmaps of row r (inserted before the last barrier).
__device__ __forceinline__ uint32_t hash_get(const RowHash &H, uint32_t r) {
    uint32_t h = hash_home(r);
    for (int p = 0; p < HASH_CAP; ++p) {
        const uint32_t k = H.key[h];
        if (k == r) return H.val[h];
        if (k == HASH_EMPTY) return 0u;
        h = (h + 1) & (HASH_CAP - 1);
    }
    return 0u;
}

// Insert the entries of nc columns (slot t / M, bit t % M) into the hash.
__device__ inline void hash_scatter(RowHash &H, int M, const ColSpan *cols, const int *pre, int nc, int total,
                                    const int32_t *crow) {
    for (int e0 = threadIdx.x; e0 < total; e0 += blockDim.x * TAB_U) {
        int t[TAB_U], r[TAB_U];
        table_entries(cols, pre, nc, total, crow, e0, blockDim.x, t, r);
#pragma unroll
        for (int u = 0; u < TAB_U; ++u)
            if (t[u] >= 0) hash_or(H, (uint32_t)r[u], (1u << (M - 1 - t[u] % M)) << (8 * (t[u] / M)));
    }
}

__device__ inline void table_scatter(uint32_t *tab, int M, int R, const ColSpan *cols, const int *pre,
                                     int nc, int total, const int32_t *crow) {
    table_pass<false>(tab, 0, M, R, cols, pre, nc, total, crow, 0, 1);
}

__device__ inline void table_clear(uint32_t *tab, int M, int R, const ColSpan *cols, const int *pre,
                                   int nc, int total, const int32_t *crow) {
    table_pass<true>(tab, 0, M, R, cols, pre, nc, total, crow, 0, 1);
}

}  // namespace vnm
