// Per-segment row tables shared by the N:M engine and the learned proposer.
//
// A table slot holds one byte per row (R bytes, 4 rows per 32-bit word): the
// M-bit bitmap of that row in one segment, MSB = the segment's first column
// (reorder/reorder.h:87-102).  A slot is filled by atomicOr from the
// segment's M column lists (CSC over run-start ids, current column s*M + b =
// run-start column inv[s*M + b]), read back once per entry, then cleared by
// the same entries, so slots are all-zero between uses.
#pragma once

#include "hip_common.h"

namespace vnm {

// Column list of slot j, bit b: run-start column inv[seg*M + b].
struct ColSpan { int start, len; };

// Fill per-column spans for K segments (K*M columns) and their prefix sums.
// Must be called by all threads; returns the total entry count.
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
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < nc; ++t) {
            pre[t] = acc;
            acc += cols[t].len;
        }
        pre[nc] = acc;
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

__device__ inline void table_scatter(uint32_t *tab, const int *slot_of_col, int M, int R,
                              const ColSpan *cols, const int *pre, int nc, int total,
                              const int32_t *crow) {
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int t = entry_col(pre, nc, e);
        int j = slot_of_col ? slot_of_col[t] : t / M;
        int b = t % M;
        int r = crow[cols[t].start + (e - pre[t])];
        atomicOr(slot_word(tab, j, R, r), (1u << (M - 1 - b)) << ((r & 3) * 8));
    }
}

__device__ inline void table_clear(uint32_t *tab, const int *slot_of_col, int M, int R,
                            const ColSpan *cols, const int *pre, int nc, int total,
                            const int32_t *crow) {
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int t = entry_col(pre, nc, e);
        int j = slot_of_col ? slot_of_col[t] : t / M;
        int r = crow[cols[t].start + (e - pre[t])];
        st_l2(slot_word(tab, j, R, r), 0u);
    }
}

}  // namespace vnm
