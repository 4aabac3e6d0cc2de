// Device-wide radix sort, scan and stream compaction for gfx950 (wave64),
// written for this library: the VENOM row sort (reordering.py:219, Python's
// stable `sorted` over per-row keys), the packer's (row, column) and
// (block, row) orders and the proposer's adjacency builds.
//
// Radix sort: LSD, 8-bit digits, one pass per digit of [0, end_bit).  A tile
// is 4,096 keys, 1,024 per wave of a 256-thread workgroup.  Per pass:
//   k_radix_hist     the tile's digit counts (LDS atomics), written
//                    digit-major (hist[d * tiles + t]);
//   scan             exclusive sum of the digit-major counts: the output
//                    offset of (digit d, tile t) -- all smaller digits first,
//                    then the earlier tiles' keys of digit d;
//   k_radix_scatter  each wave ranks its 1,024 keys in 16 rounds of 64: the
//                    lanes of a round holding one digit find each other with
//                    8 ballots (one per digit bit), the lowest of them adds
//                    the group's size to the wave's running count of the
//                    digit (LDS), so a key's rank is its wave's count so far
//                    plus the lanes below it in the group.  Waves' counts are
//                    then prefixed in wave order, the tile's keys placed in
//                    LDS in digit order, and written out with consecutive
//                    lanes on consecutive output positions.
// Every step keeps input order among equal digits, so each pass -- and the
// sort -- is stable.  Passes ping-pong between the output and a scratch
// buffer so the last one lands in the output; the input is never written.
//
// Scan: tiles of 4,096 elements; per-tile totals (k_scan_reduce), one
// workgroup scanning the totals (k_scan_totals), then each tile scanned in
// LDS with its offset (k_scan_tile).  Sum and max, int32 or uint8 input.
// Inputs of at most one tile are sorted (k_radix_small) or scanned
// (k_scan_small) by one workgroup in one launch.
// Compaction: exclusive sum of the flags, then a scatter of flagged items.
#include "sortlib.h"

#include <algorithm>

namespace vnm {

namespace {

constexpr int WG = 256;                 // threads per workgroup
constexpr int NWAVE = WG / 64;
constexpr int ROUNDS = 16;              // 64-key rounds per wave and tile
constexpr int TILE = WG * ROUNDS;       // 4,096 keys per tile
constexpr int RADIX = 256;

// Scratch of the calling host thread, grown on demand (one per thread, so
// concurrent contexts on worker threads do not share it).
struct TempBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes && p) return VNM_OK;
        if (p) (void)hipFree(p);  // waits for the device, so no kernel still reads it
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, need ? need : 1) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(VNM_E_OOM, "scratch allocation of %zu bytes failed", need);
        }
        bytes = need;
        return VNM_OK;
    }
    ~TempBuf() {
        if (p) (void)hipFree(p);
    }
};
// Scratch slots per host thread; the current one serves every sort and scan
// (separate slots let launch chains that may run concurrently, such as the
// branches of one captured graph, sort at the same time).
constexpr int SCRATCH_SLOTS = 16;
thread_local TempBuf g_slots[SCRATCH_SLOTS];
thread_local int g_slot = 0;
#define g_tmp (g_slots[g_slot])

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Inclusive sum over the 64 lanes of a wave.
template <typename T, typename Op>
__device__ __forceinline__ T wave_scan_incl(T v, Op op) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = __shfl_up(v, d, 64);
        if (lane >= d) v = op(v, o);
    }
    return v;
}

// Exclusive scan of one value per thread over the workgroup (wave order),
// identity `id`; *total = the workgroup's reduction.  Uses `sh` (NWAVE Ts).
template <typename T, typename Op>
__device__ __forceinline__ T wg_scan_excl(T v, Op op, T id, T *sh, T *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T inc = wave_scan_incl(v, op);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    T base = id, all = id;
#pragma unroll
    for (int i = 0; i < NWAVE; ++i) {
        if (i < w) base = op(base, sh[i]);
        all = op(all, sh[i]);
    }
    __syncthreads();
    *total = all;
    T ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = id;
    return op(base, ex);
}

struct OpSum {
    __device__ __forceinline__ int32_t operator()(int32_t a, int32_t b) const { return a + b; }
    static constexpr int32_t id = 0;
};
struct OpMax {
    __device__ __forceinline__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
    static constexpr int32_t id = INT32_MIN;
};

// ------------------------------------------------------------ radix sort

__device__ __forceinline__ int digit_of(uint64_t k, int shift, int bits) {
    return (int)((k >> shift) & ((1u << bits) - 1u));
}

__global__ __launch_bounds__(WG) void k_radix_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                   int bits, int tiles, int32_t *__restrict__ hist, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t cnt[RADIX];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * TILE;
#pragma unroll 4
    for (int j = 0; j < ROUNDS; ++j) {
        const int64_t i = t0 + j * WG + threadIdx.x;
        if (i < n) atomicAdd(&cnt[digit_of(keys[i], shift, bits)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

template <bool VALS>
__global__ __launch_bounds__(WG) void k_radix_scatter(const uint64_t *__restrict__ kin,
                                                      const int32_t *__restrict__ vin,
                                                      uint64_t *__restrict__ kout, int32_t *__restrict__ vout,
                                                      int64_t n, int shift, int bits, int tiles,
                                                      const int32_t *__restrict__ offs, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t run[NWAVE][RADIX];   // per wave: keys of each digit so far
    __shared__ int32_t start[RADIX];        // tile position of each digit's first key
    __shared__ int32_t sh[NWAVE];
    __shared__ uint64_t sk[TILE];
    __shared__ int32_t sv[VALS ? TILE : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < NWAVE * RADIX; i += WG) (&run[0][0])[i] = 0;
    __syncthreads();
    // wave w ranks keys [w * 1024, (w + 1) * 1024) of the tile, 64 per round
    const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)w * (64 * ROUNDS);
    uint64_t key[ROUNDS];
    int32_t val[ROUNDS];
    int dig[ROUNDS], rnk[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int64_t i = base + r * 64 + lane;
        const bool ok = i < n;
        key[r] = ok ? kin[i] : 0ull;
        if constexpr (VALS) val[r] = ok ? vin[i] : 0;
        const int d = ok ? digit_of(key[r], shift, bits) : 0;
        uint64_t peers = __ballot(ok);
        for (int b = 0; b < bits; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        const int below = __popcll(peers & lanemask_lt());
        int r0 = 0;
        if (ok) r0 = run[w][d];
        // the group's lowest lane advances the digit's count after every
        // lane of the group has read it (same wave: LDS order is program order)
        if (ok && below == 0) run[w][d] = r0 + __popcll(peers);
        dig[r] = ok ? d : -1;
        rnk[r] = r0 + below;
    }
    __syncthreads();
    // per digit (thread d): prefix over waves in wave order, tile count
    const int d = threadIdx.x;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < NWAVE; ++i) {
        const int32_t x = run[i][d];
        run[i][d] = c;
        c += x;
    }
    int32_t total;
    const int32_t st = wg_scan_excl<int32_t>(c, OpSum(), 0, sh, &total);
    start[d] = st;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        if (dig[r] >= 0) {
            const int p = start[dig[r]] + run[w][dig[r]] + rnk[r];
            sk[p] = key[r];
            if constexpr (VALS) sv[p] = val[r];
        }
    }
    __syncthreads();
    // write out: tile position p holds the key of digit dd at offset
    // p - start[dd] of that digit's run
    const int64_t t0 = (int64_t)blockIdx.x * TILE;
    const int cnt = (int)std::min<int64_t>(TILE, n - t0);
    for (int p = threadIdx.x; p < cnt; p += WG) {
        const uint64_t k = sk[p];
        const int dd = digit_of(k, shift, bits);
        const int64_t o = (int64_t)offs[(int64_t)dd * tiles + blockIdx.x] + (p - start[dd]);
        kout[o] = k;
        if constexpr (VALS) vout[o] = sv[p];
    }
}

// A whole sort of at most TILE keys in one workgroup: every pass ranks the
// keys as k_radix_scatter does (each wave its 1,024 keys in 16 rounds of
// 64), places them in LDS in digit order and reads them back in the rank
// layout for the next pass -- one launch instead of five per pass, for the
// many small sorts of the VENOM refinement levels and the proposer.
template <bool VALS>
__global__ __launch_bounds__(WG) void k_radix_small(const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin,
                                                    uint64_t *__restrict__ kout, int32_t *__restrict__ vout, int n,
                                                    int end_bit, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t run[NWAVE][RADIX];
    __shared__ int32_t start[RADIX];
    __shared__ int32_t sh[NWAVE];
    __shared__ uint64_t sk[TILE];
    __shared__ int32_t sv[VALS ? TILE : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int base = w * (64 * ROUNDS);
    uint64_t key[ROUNDS];
    int32_t val[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int i = base + r * 64 + lane;
        key[r] = i < n ? kin[i] : 0ull;
        if constexpr (VALS) val[r] = i < n ? vin[i] : 0;
    }
    for (int shift = 0; shift < end_bit; shift += 8) {
        const int bits = min(8, end_bit - shift);
        for (int i = threadIdx.x; i < NWAVE * RADIX; i += WG) (&run[0][0])[i] = 0;
        __syncthreads();
        int dig[ROUNDS], rnk[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const bool ok = base + r * 64 + lane < n;
            const int d = ok ? digit_of(key[r], shift, bits) : 0;
            uint64_t peers = __ballot(ok);
            for (int b = 0; b < bits; ++b) {
                const uint64_t bal = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? bal : ~bal;
            }
            const int below = __popcll(peers & lanemask_lt());
            int r0 = 0;
            if (ok) r0 = run[w][d];
            if (ok && below == 0) run[w][d] = r0 + __popcll(peers);
            dig[r] = ok ? d : -1;
            rnk[r] = r0 + below;
        }
        __syncthreads();
        const int d = threadIdx.x;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < NWAVE; ++i) {
            const int32_t x = run[i][d];
            run[i][d] = c;
            c += x;
        }
        int32_t total;
        start[d] = wg_scan_excl<int32_t>(c, OpSum(), 0, sh, &total);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            if (dig[r] >= 0) {
                const int p = start[dig[r]] + run[w][dig[r]] + rnk[r];
                sk[p] = key[r];
                if constexpr (VALS) sv[p] = val[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const int i = base + r * 64 + lane;
            if (i < n) {
                key[r] = sk[i];
                if constexpr (VALS) val[r] = sv[i];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int i = base + r * 64 + lane;
        if (i < n) {
            kout[i] = key[r];
            if constexpr (VALS) vout[i] = val[r];
        }
    }
}

// Sorts of 2 to FUSED_TILES tiles: one launch per pass.  Each workgroup
// derives its tile's offsets itself from the pass's (digit, tile) count
// table (a scan of at most 256 * FUSED_TILES counts, in (digit, tile)
// order), ranks and scatters its tile as k_radix_scatter does, and counts
// the next pass's digits of the keys it writes by destination tile (an LDS
// table, then one atomic add per nonzero count), so no separate histogram
// or scan launch is needed between passes.
constexpr int FUSED_TILES = 16;

__global__ void k_fill_zero_i32(int32_t *p, int n, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <bool VALS>
__global__ __launch_bounds__(WG) void k_radix_fused(const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin,
                                                    uint64_t *__restrict__ kout, int32_t *__restrict__ vout,
                                                    int n, int shift, int bits, int tiles,
                                                    const int32_t *__restrict__ hist, int next_shift,
                                                    int next_bits, int32_t *__restrict__ next_hist, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t run[NWAVE][RADIX];
    __shared__ int32_t start[RADIX];
    __shared__ int32_t offs[RADIX];
    __shared__ int32_t sh[NWAVE];
    __shared__ uint64_t sk[TILE];
    __shared__ int32_t sv[VALS ? TILE : 1];
    __shared__ int32_t nh[FUSED_TILES * RADIX];  // next pass: counts per (destination tile, digit)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = blockIdx.x;
    for (int i = threadIdx.x; i < NWAVE * RADIX; i += WG) (&run[0][0])[i] = 0;
    if (next_hist)
        for (int i = threadIdx.x; i < tiles * RADIX; i += WG) nh[i] = 0;
    // this tile's offsets: digits before d over all tiles, then digit d in
    // the tiles before this one (hist is digit-major: hist[d * tiles + t])
    {
        const int d = threadIdx.x;
        int32_t row = 0, before = 0;
        for (int t = 0; t < tiles; ++t) {
            const int32_t x = hist[d * tiles + t];
            before += t < tile ? x : 0;
            row += x;
        }
        int32_t total;
        offs[d] = wg_scan_excl<int32_t>(row, OpSum(), 0, sh, &total) + before;
    }
    __syncthreads();
    const int base = tile * TILE + w * (64 * ROUNDS);
    uint64_t key[ROUNDS];
    int32_t val[ROUNDS];
    int dig[ROUNDS], rnk[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int i = base + r * 64 + lane;
        const bool ok = i < n;
        key[r] = ok ? kin[i] : 0ull;
        if constexpr (VALS) val[r] = ok ? vin[i] : 0;
        const int d = ok ? digit_of(key[r], shift, bits) : 0;
        uint64_t peers = __ballot(ok);
        for (int b = 0; b < bits; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        const int below = __popcll(peers & lanemask_lt());
        int r0 = 0;
        if (ok) r0 = run[w][d];
        if (ok && below == 0) run[w][d] = r0 + __popcll(peers);
        dig[r] = ok ? d : -1;
        rnk[r] = r0 + below;
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < NWAVE; ++i) {
            const int32_t x = run[i][d];
            run[i][d] = c;
            c += x;
        }
        int32_t total;
        start[d] = wg_scan_excl<int32_t>(c, OpSum(), 0, sh, &total);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        if (dig[r] >= 0) {
            const int p = start[dig[r]] + run[w][dig[r]] + rnk[r];
            sk[p] = key[r];
            if constexpr (VALS) sv[p] = val[r];
        }
    }
    __syncthreads();
    const int cnt = min(TILE, n - tile * TILE);
    for (int p = threadIdx.x; p < cnt; p += WG) {
        const uint64_t k = sk[p];
        const int dd = digit_of(k, shift, bits);
        const int o = offs[dd] + (p - start[dd]);
        kout[o] = k;
        if constexpr (VALS) vout[o] = sv[p];
        if (next_hist) atomicAdd(&nh[(o / TILE) * RADIX + digit_of(k, next_shift, next_bits)], 1);
    }
    if (next_hist) {
        __syncthreads();
        for (int i = threadIdx.x; i < tiles * RADIX; i += WG) {
            const int32_t c = nh[i];
            if (c) atomicAdd(&next_hist[(i % RADIX) * tiles + i / RADIX], c);
        }
    }
}

// ------------------------------------------------------------ scan

// Scan input as int32: flags count as 0 or 1.
__device__ __forceinline__ int32_t load_as(const int32_t *p, int64_t i) { return p[i]; }
__device__ __forceinline__ int32_t load_as(const uint8_t *p, int64_t i) { return p[i] != 0; }

template <typename In, typename Op>
__global__ __launch_bounds__(WG) void k_scan_reduce(const In *__restrict__ in, int64_t n,
                                                    int32_t *__restrict__ totals, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t sh[NWAVE];
    const Op op;
    const int64_t t0 = (int64_t)blockIdx.x * TILE;
    int32_t acc = Op::id;
    for (int j = 0; j < ROUNDS; ++j) {
        const int64_t i = t0 + j * WG + threadIdx.x;
        if (i < n) acc = op(acc, load_as(in, i));
    }
    int32_t total;
    (void)wg_scan_excl<int32_t>(acc, op, Op::id, sh, &total);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the tile totals in place.
template <typename Op>
__global__ __launch_bounds__(WG) void k_scan_totals(int32_t *__restrict__ totals, int tiles, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t sh[NWAVE];
    const Op op;
    int32_t carry = Op::id;
    for (int b = 0; b < tiles; b += WG) {
        const int i = b + threadIdx.x;
        const int32_t v = i < tiles ? totals[i] : Op::id;
        int32_t total;
        const int32_t ex = wg_scan_excl<int32_t>(v, op, Op::id, sh, &total);
        if (i < tiles) totals[i] = op(carry, ex);
        carry = op(carry, total);
    }
}

// Tile scan: elements staged in LDS, each thread scans 16 consecutive ones.
template <typename In, typename Op, bool EXCL>
__global__ __launch_bounds__(WG) void k_scan_tile(const In *__restrict__ in, int32_t *__restrict__ out, int64_t n,
                                                  const int32_t *__restrict__ offsets, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t s[TILE + TILE / 32];  // one pad word per 32 to spread the banks
    __shared__ int32_t sh[NWAVE];
    const Op op;
    const int64_t t0 = (int64_t)blockIdx.x * TILE;
    auto at = [](int i) { return i + (i >> 5); };
    for (int j = 0; j < ROUNDS; ++j) {
        const int i = j * WG + threadIdx.x;
        s[at(i)] = t0 + i < n ? load_as(in, t0 + i) : Op::id;
    }
    __syncthreads();
    int32_t v[ROUNDS];
    int32_t acc = Op::id;
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
        v[j] = s[at(threadIdx.x * ROUNDS + j)];
        acc = op(acc, v[j]);
    }
    int32_t total;
    int32_t run = op(offsets[blockIdx.x], wg_scan_excl<int32_t>(acc, op, Op::id, sh, &total));
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
        const int32_t next = op(run, v[j]);
        s[at(threadIdx.x * ROUNDS + j)] = EXCL ? run : next;
        run = next;
    }
    __syncthreads();
    for (int j = 0; j < ROUNDS; ++j) {
        const int i = j * WG + threadIdx.x;
        if (t0 + i < n) out[t0 + i] = s[at(i)];
    }
}

template <typename In, typename Op, bool EXCL>
// A whole scan of at most TILE elements in one workgroup (one launch).
__global__ __launch_bounds__(WG) void k_scan_small(const In *__restrict__ in, int32_t *__restrict__ out, int n, const int *__restrict__ skip) {
    if (skip && *skip) return;  // the caller's flag: this sort is not wanted
    __shared__ int32_t s[TILE + TILE / 32];
    __shared__ int32_t sh[NWAVE];
    const Op op;
    auto at = [](int i) { return i + (i >> 5); };
    for (int j = 0; j < ROUNDS; ++j) {
        const int i = j * WG + threadIdx.x;
        s[at(i)] = i < n ? load_as(in, i) : Op::id;
    }
    __syncthreads();
    int32_t v[ROUNDS];
    int32_t acc = Op::id;
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
        v[j] = s[at(threadIdx.x * ROUNDS + j)];
        acc = op(acc, v[j]);
    }
    int32_t total;
    int32_t run = wg_scan_excl<int32_t>(acc, op, Op::id, sh, &total);
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
        const int32_t next = op(run, v[j]);
        s[at(threadIdx.x * ROUNDS + j)] = EXCL ? run : next;
        run = next;
    }
    __syncthreads();
    for (int j = 0; j < ROUNDS; ++j) {
        const int i = j * WG + threadIdx.x;
        if (i < n) out[i] = s[at(i)];
    }
}

__global__ void k_select_scatter(const int32_t *__restrict__ in, const uint8_t *__restrict__ flags,
                                 const int32_t *__restrict__ pos, int64_t n, int32_t *__restrict__ out,
                                 int32_t *__restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool f = flags[i] != 0;
    if (f) out[pos[i]] = in[i];
    if (i == n - 1) *count = pos[i] + (f ? 1 : 0);
}

int tiles_of(int64_t n) { return (int)((n + TILE - 1) / TILE); }

// Scan of n elements with `totals` (>= tiles_of(n) ints) as scratch.
template <typename In, typename Op, bool EXCL>
int scan_with(const In *in, int32_t *out, int64_t n, int32_t *totals, hipStream_t st,
              const int *skip = nullptr) {
    if (n <= TILE) {
        hipLaunchKernelGGL((k_scan_small<In, Op, EXCL>), dim3(1), dim3(WG), 0, st, in, out, (int)n, skip);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(VNM_E_HIP, "scan launch failed: %s", hipGetErrorString(e));
        return VNM_OK;
    }
    const int tiles = tiles_of(n);
    hipLaunchKernelGGL((k_scan_reduce<In, Op>), dim3(tiles), dim3(WG), 0, st, in, n, totals, skip);
    hipLaunchKernelGGL((k_scan_totals<Op>), dim3(1), dim3(WG), 0, st, totals, tiles, skip);
    hipLaunchKernelGGL((k_scan_tile<In, Op, EXCL>), dim3(tiles), dim3(WG), 0, st, in, out, n, totals, skip);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(VNM_E_HIP, "scan launch failed: %s", hipGetErrorString(e));
    return VNM_OK;
}

template <bool VALS>
int radix_sort(const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout, int64_t n, int end_bit,
               hipStream_t st, const int *skip = nullptr) {
    if (n <= 0) return VNM_OK;
    if (n > (int64_t)INT32_MAX) return set_error(VNM_E_ARG, "radix sort of %lld keys", (long long)n);
    end_bit = std::max(1, std::min(end_bit, 64));
    if (n <= TILE) {
        hipLaunchKernelGGL((k_radix_small<VALS>), dim3(1), dim3(WG), 0, st, kin, vin, kout, vout, (int)n, end_bit, skip);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(VNM_E_HIP, "radix sort launch failed: %s", hipGetErrorString(e));
        return VNM_OK;
    }
    const int passes = (end_bit + 7) / 8;
    const int tiles = tiles_of(n);
    if (tiles <= FUSED_TILES) {
        // one count table per pass, zeroed up front; pass 0's by k_radix_hist,
        // each later one by the pass before it
        const size_t kb = align256((size_t)n * 8), vb = VALS ? align256((size_t)n * 4) : 0;
        const int tab = RADIX * tiles;
        int rc = g_tmp.ensure(kb + vb + align256((size_t)tab * passes * 4));
        if (rc) return rc;
        uint64_t *ktmp = (uint64_t *)g_tmp.p;
        int32_t *vtmp = (int32_t *)((char *)g_tmp.p + kb);
        int32_t *hists = (int32_t *)((char *)g_tmp.p + kb + vb);
        hipLaunchKernelGGL(k_fill_zero_i32, dim3((tab * passes + 255) / 256), dim3(256), 0, st, hists, tab * passes, skip);
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(WG), 0, st, kin, n, 0, std::min(8, end_bit), tiles, hists, skip);
        const uint64_t *ks = kin;
        const int32_t *vs = vin;
        for (int p = 0; p < passes; ++p) {
            const int shift = 8 * p, bits = std::min(8, end_bit - shift);
            const bool last = p + 1 == passes;
            uint64_t *kd = ((passes - 1 - p) % 2 == 0) ? kout : ktmp;
            int32_t *vd = ((passes - 1 - p) % 2 == 0) ? vout : vtmp;
            hipLaunchKernelGGL((k_radix_fused<VALS>), dim3(tiles), dim3(WG), 0, st, ks, vs, kd, vd, (int)n, shift, bits,
                               tiles, hists + (size_t)tab * p, last ? 0 : shift + 8,
                               last ? 1 : std::min(8, end_bit - shift - 8), last ? nullptr : hists + (size_t)tab * (p + 1), skip);
            ks = kd;
            vs = vd;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(VNM_E_HIP, "radix sort launch failed: %s", hipGetErrorString(e));
        return VNM_OK;
    }
    const int64_t nh = (int64_t)RADIX * tiles;
    // scratch: keys and values of the other buffer, digit counts, offsets,
    // scan totals
    const size_t kb = align256((size_t)n * 8), vb = VALS ? align256((size_t)n * 4) : 0;
    const size_t hb = align256((size_t)nh * 4), tb = align256((size_t)tiles_of(nh) * 4);
    int rc = g_tmp.ensure(kb + vb + 2 * hb + tb);
    if (rc) return rc;
    char *base = (char *)g_tmp.p;
    uint64_t *ktmp = (uint64_t *)base;
    int32_t *vtmp = (int32_t *)(base + kb);
    int32_t *hist = (int32_t *)(base + kb + vb);
    int32_t *offs = (int32_t *)(base + kb + vb + hb);
    int32_t *tot = (int32_t *)(base + kb + vb + 2 * hb);
    const uint64_t *ks = kin;
    const int32_t *vs = vin;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p, bits = std::min(8, end_bit - shift);
        // the last pass writes the output
        uint64_t *kd = ((passes - 1 - p) % 2 == 0) ? kout : ktmp;
        int32_t *vd = ((passes - 1 - p) % 2 == 0) ? vout : vtmp;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(WG), 0, st, ks, n, shift, bits, tiles, hist, skip);
        if ((rc = scan_with<int32_t, OpSum, true>(hist, offs, nh, tot, st, skip))) return rc;
        hipLaunchKernelGGL((k_radix_scatter<VALS>), dim3(tiles), dim3(WG), 0, st, ks, vs, kd, vd, n, shift, bits,
                           tiles, offs, skip);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(VNM_E_HIP, "radix sort launch failed: %s", hipGetErrorString(e));
        ks = kd;
        vs = vd;
    }
    return VNM_OK;
}

}  // namespace

int sort_keys_u64(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t st) {
    return radix_sort<false>(in, out, nullptr, nullptr, n, end_bit, st);
}

int sort_keys_u64_if(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t st, const int *skip) {
    return radix_sort<false>(in, out, nullptr, nullptr, n, end_bit, st, skip);
}

int sort_scratch_slot(int slot) {
    const int prev = g_slot;
    if (slot >= 0 && slot < SCRATCH_SLOTS) g_slot = slot;
    return prev;
}

int sort_scratch_reserve(int64_t n) {
    if (n <= TILE) return VNM_OK;
    // the larger of the two multi-tile paths' needs, keys and values, 8 passes
    const int tiles = tiles_of(n);
    const size_t kb = align256((size_t)n * 8), vb = align256((size_t)n * 4);
    const size_t fused = kb + vb + align256((size_t)RADIX * tiles * 8 * 4);
    const int64_t nh = (int64_t)RADIX * tiles;
    const size_t multi = kb + vb + 2 * align256((size_t)nh * 4) + align256((size_t)tiles_of(nh) * 4);
    return g_tmp.ensure(std::max(fused, multi));
}

int sort_pairs_u64_i32(const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout, int64_t n,
                       int end_bit, hipStream_t st) {
    return radix_sort<true>(kin, kout, vin, vout, n, end_bit, st);
}

int scan_excl_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t st) {
    if (n <= 0) return VNM_OK;
    int rc = g_tmp.ensure(align256((size_t)tiles_of(n) * 4));
    if (rc) return rc;
    return scan_with<int32_t, OpSum, true>(in, out, n, (int32_t *)g_tmp.p, st);
}

int scan_incl_max_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t st) {
    if (n <= 0) return VNM_OK;
    int rc = g_tmp.ensure(align256((size_t)tiles_of(n) * 4));
    if (rc) return rc;
    return scan_with<int32_t, OpMax, false>(in, out, n, (int32_t *)g_tmp.p, st);
}

int select_flagged_i32(const int32_t *in, const uint8_t *flags, int32_t *out, int32_t *d_count, int64_t n,
                       hipStream_t st) {
    if (n <= 0) {
        hipError_t e = hipMemsetAsync(d_count, 0, sizeof(int32_t), st);
        return e == hipSuccess ? VNM_OK : set_error(VNM_E_HIP, "select: %s", hipGetErrorString(e));
    }
    const size_t pb = align256((size_t)n * 4);
    int rc = g_tmp.ensure(pb + align256((size_t)tiles_of(n) * 4));
    if (rc) return rc;
    int32_t *pos = (int32_t *)g_tmp.p;
    if ((rc = scan_with<uint8_t, OpSum, true>(flags, pos, n, (int32_t *)((char *)g_tmp.p + pb), st))) return rc;
    hipLaunchKernelGGL(k_select_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, flags, pos, n, out,
                       d_count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? VNM_OK : set_error(VNM_E_HIP, "select launch failed: %s", hipGetErrorString(e));
}

}  // namespace vnm

#include "graph.h"

extern "C" int vnm_sort_pairs(vnm_ctx *ctx, const uint64_t *keys_in, uint64_t *keys_out, const int32_t *vals_in,
                              int32_t *vals_out, int64_t n, int end_bit) {
    if (!ctx || (n > 0 && (!keys_in || !keys_out))) return vnm::set_error(VNM_E_ARG, "sort: null argument");
    if (n < 0 || end_bit < 1 || end_bit > 64) return vnm::set_error(VNM_E_ARG, "sort: n >= 0, end_bit in [1, 64]");
    if ((vals_in == nullptr) != (vals_out == nullptr))
        return vnm::set_error(VNM_E_ARG, "sort: values in and out must both be given or both be NULL");
    if (keys_in == keys_out || (vals_in && vals_in == vals_out))
        return vnm::set_error(VNM_E_ARG, "sort: output must not alias the input");
    int rc = vals_in ? vnm::sort_pairs_u64_i32(keys_in, keys_out, vals_in, vals_out, n, end_bit, ctx->stream)
                     : vnm::sort_keys_u64(keys_in, keys_out, n, end_bit, ctx->stream);
    if (rc) return rc;
    ctx->launches += 5 * ((end_bit + 7) / 8);
    return VNM_OK;
}
