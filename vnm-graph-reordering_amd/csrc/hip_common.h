// Device-side helpers shared by the HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdint.h>

#include "vnm_internal.h"

namespace vnm {

#define VNM_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return set_error(VNM_E_HIP, "%s failed: %s (%s:%d)", #expr,                \
                             hipGetErrorString(_e), __FILE__, __LINE__);               \
    } while (0)

#define VNM_LAUNCH_CHECK(name)                                                         \
    do {                                                                               \
        hipError_t _e = hipGetLastError();                                             \
        if (_e != hipSuccess)                                                          \
            return set_error(VNM_E_HIP, "launch of %s failed: %s", name,               \
                             hipGetErrorString(_e));                                   \
    } while (0)

// Loads/stores of memory another thread of the same workgroup (or a later
// kernel phase) updated through L2 atomics: agent-scope relaxed accesses skip
// the non-coherent vector L1.
template <typename T>
__device__ __forceinline__ T ld_l2(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_l2(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Priority key of a SwapPair under the reference Comparator
// (reorder/swappair.h:112-120): Comparator(p1, p2) == key(p1) <= key(p2).
// The gain sum dominates; equal sums compare right gains.
__device__ __forceinline__ int64_t swap_key(int lg, int rg) {
    return ((int64_t)(lg + rg) << 32) | (int64_t)(uint32_t)(rg ^ (int)0x80000000);
}

// std::priority_queue<SwapPair*> with the non-strict Comparator, emulated
// exactly as libstdc++ pushes (push_back + __push_heap) and pops (pop_heap ->
// __adjust_heap -> __push_heap).  A heap entry packs rank << 8 | element,
// where rank is the element's dense rank among the list's keys (the number
// of strictly smaller keys), so Comparator(a, b) holds exactly when
// rank(a) <= rank(b) and each comparison reads one entry.  One thread owns a
// heap; slot s sits at h[s * stride], so the threads of a wave working on
// interleaved heaps touch consecutive LDS words.
struct HeapRef {
    uint16_t *h;
    int n;       // current size
    int stride;
};

__device__ __forceinline__ uint32_t hget(const HeapRef &H, int slot) { return H.h[slot * H.stride]; }
__device__ __forceinline__ void hset(HeapRef &H, int slot, uint32_t v) { H.h[slot * H.stride] = (uint16_t)v; }
__device__ __forceinline__ bool heap_le(uint32_t a, uint32_t b) { return (a >> 8) <= (b >> 8); }

constexpr int HEAP_DEPTH = 6;  // heaps of at most 64 entries (M <= 8)

// __push_heap(first, hole, top, value): the hole moves up while the parent
// p satisfies Comparator(p, value).  The path hole, (hole - 1) / 2, ... is
// fixed, so every ancestor on it is read at once (independent LDS reads
// instead of one dependent read per level); the passed ancestors then move
// down one slot and the value takes the last hole.  The slots read are above
// every slot written, so reading first changes nothing.
__device__ __forceinline__ void heap_sift_up(HeapRef &H, int hole, int top, uint32_t value) {
    int pos[HEAP_DEPTH + 1];
    uint32_t anc[HEAP_DEPTH];
    pos[0] = hole;
#pragma unroll
    for (int j = 0; j < HEAP_DEPTH; ++j) {
        const bool up = pos[j] > top;
        pos[j + 1] = up ? (pos[j] - 1) / 2 : pos[j];
        anc[j] = up ? hget(H, pos[j + 1]) : 0u;
    }
    int t = 0;
    bool go = true;
#pragma unroll
    for (int j = 0; j < HEAP_DEPTH; ++j) {
        go = go && pos[j] > top && heap_le(anc[j], value);
        if (go) {
            hset(H, pos[j], anc[j]);
            t = j + 1;
        }
    }
    hset(H, pos[t], value);
}

__device__ __forceinline__ void heap_push(HeapRef &H, uint32_t value) {
    H.n += 1;
    heap_sift_up(H, H.n - 1, 0, value);
}

// Returns the popped entry (the former top).
__device__ __forceinline__ uint32_t heap_pop(HeapRef &H) {
    const uint32_t top_el = hget(H, 0);
    const int n = H.n;
    if (n > 1) {
        const int last = n - 1;
        const uint32_t value = hget(H, last);
        hset(H, last, top_el);
        // __adjust_heap(first, 0, len = last, value)
        int len = last, hole = 0, second = 0;
        while (second < (len - 1) / 2) {
            second = 2 * (second + 1);
            const uint32_t r = hget(H, second), l = hget(H, second - 1);
            uint32_t c = r;
            if (heap_le(r, l)) { second--; c = l; }
            hset(H, hole, c);
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            hset(H, hole, hget(H, second - 1));
            hole = second - 1;
        }
        heap_sift_up(H, hole, 0, value);
    }
    H.n = n - 1;
    return top_el;
}

// BF16 storage codec: round to nearest even, NaN kept quiet.
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ int popc32(uint32_t x) { return __builtin_popcount(x); }

}  // namespace vnm
