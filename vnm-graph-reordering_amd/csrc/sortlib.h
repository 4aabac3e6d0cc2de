// Device-wide sort/scan/select primitives used by the VENOM, packing and
// proposer passes (hand-written in sortlib.hip).  Scratch is per host thread
// and kernels run on `st`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vnm_internal.h"

namespace vnm {

// Stable LSD radix sort on bits [0, end_bit) of the keys (the input is not
// written; out must not alias it).
int sort_keys_u64(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t st);
// sort_keys_u64 whose kernels all return at once while *skip is nonzero
// (read on the device when they run: a launch chain enqueued ahead of the
// decision)
int sort_keys_u64_if(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t st, const int *skip);
// Grow this host thread's current scratch slot to what a sort of n keys
// (with values) needs, so that sorts up to that size allocate nothing (as
// under capture).
int sort_scratch_reserve(int64_t n);
// Make scratch slot `slot` (0..15) of this host thread the current one;
// returns the previous slot.  Sorts enqueued on streams that may run
// concurrently need separate slots.
int sort_scratch_slot(int slot);
int sort_pairs_u64_i32(const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout,
                       int64_t n, int end_bit, hipStream_t st);
int scan_excl_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t st);
int scan_incl_max_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t st);
// out[j] = in[i] for the j-th i with flags[i] != 0, in order; *d_count = j's.
int select_flagged_i32(const int32_t *in, const uint8_t *flags, int32_t *out, int32_t *d_count,
                       int64_t n, hipStream_t st);

}  // namespace vnm
