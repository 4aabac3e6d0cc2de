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
// heap of at most 64 entries in 33 words of its own: slot s (0-based) is the
// 16-bit entry h[s + 1], so the two children of slot s (slots 2s + 1 and
// 2s + 2, entries 2i and 2i + 1 for i = s + 1) are the low and high halves
// of word i and one read fetches both.  The odd word count puts the heaps of
// a wave's threads on different LDS banks.  Below, positions are 1-based
// (i = slot + 1): the parent of i is i / 2, the root is 1.
constexpr int HEAP_DEPTH = 6;   // heaps of at most 64 entries (M <= 8)
constexpr int HEAP_WORDS = 33;  // words per heap (entries 1..64; entry 0 and 65 unused)

struct HeapRef {
    uint16_t *h;  // entry i at h[i]; 4-byte aligned
    int n;        // current size
};

__host__ __device__ __forceinline__ uint32_t hget(const HeapRef &H, int slot) { return H.h[slot + 1]; }
__host__ __device__ __forceinline__ void hset(HeapRef &H, int slot, uint32_t v) { H.h[slot + 1] = (uint16_t)v; }
__host__ __device__ __forceinline__ uint32_t hword(const HeapRef &H, int i) {
    return reinterpret_cast<const uint32_t *>(H.h)[i];  // entries 2i, 2i + 1
}
__host__ __device__ __forceinline__ bool heap_le(uint32_t a, uint32_t b) { return (a >> 8) <= (b >> 8); }

// __push_heap(first, hole, 0, value) at 1-based position `i`: the hole moves
// up while the parent p satisfies Comparator(p, value).  The path i, i / 2,
// ... is fixed (it stays at the root once there), so every ancestor on it is
// read at once; the passed ancestors then move down one position and the
// value takes the last hole.  Every position of the path is written, from
// the top down, so the writes need no branches: with go[j] true for the
// passed levels j < t, position ip[j] gets its parent's entry anc[j] for
// j < t, the value for j == t and its own entry anc[j - 1] back for j > t;
// the root, repeated at the path's top, ends with its first occurrence's
// entry.
__host__ __device__ __forceinline__ void heap_sift_up1(HeapRef &H, int i, uint32_t value) {
    int ip[HEAP_DEPTH + 1];
    uint32_t anc[HEAP_DEPTH];
    ip[0] = i;
#pragma unroll
    for (int j = 0; j < HEAP_DEPTH; ++j) ip[j + 1] = (ip[j] >> 1) > 1 ? ip[j] >> 1 : 1;
#pragma unroll
    for (int j = 0; j < HEAP_DEPTH; ++j) anc[j] = H.h[ip[j + 1]];
    bool go[HEAP_DEPTH];
#pragma unroll
    for (int j = 0; j < HEAP_DEPTH; ++j) go[j] = (j == 0 || go[j - 1]) && ip[j] > 1 && heap_le(anc[j], value);
    H.h[ip[HEAP_DEPTH]] = (uint16_t)(go[HEAP_DEPTH - 1] ? value : anc[HEAP_DEPTH - 1]);
#pragma unroll
    for (int j = HEAP_DEPTH - 1; j >= 1; --j) H.h[ip[j]] = (uint16_t)(go[j] ? anc[j] : go[j - 1] ? value : anc[j - 1]);
    H.h[ip[0]] = (uint16_t)(go[0] ? anc[0] : value);
}

__host__ __device__ __forceinline__ void heap_push(HeapRef &H, uint32_t value) {
    H.n += 1;
    heap_sift_up1(H, H.n, value);
}

// Returns the popped entry (the former top).  The descent (__adjust_heap)
// takes two levels per step: the hole's children word and both children's
// children words are read together, so a 64-entry heap is three round trips
// deep.  The first step's words are read before the popped entry is stored
// at the last position: the descent only uses positions before it.
__host__ __device__ __forceinline__ uint32_t heap_pop(HeapRef &H) {
    const int n = H.n;
    const uint32_t top_el = H.h[1];
    if (n > 1) {
        const int len = n - 1;          // entries left after the pop
        const int lim = (len - 1) / 2;  // the descent steps while the hole is at most lim
        uint32_t cw = hword(H, 1), g0 = hword(H, 2), g1 = hword(H, 3);
        const uint32_t value = H.h[n];
        H.h[n] = (uint16_t)top_el;
        int i = 1;
        while (i <= lim) {
            // cw: the hole's children 2i, 2i + 1; g0 / g1: theirs
            const uint32_t l = cw & 0xFFFFu, r = cw >> 16;
            const bool left = heap_le(r, l);  // ties go to the left child
            const int c1 = 2 * i + (left ? 0 : 1);
            H.h[i] = (uint16_t)(left ? l : r);
            i = c1;
            if (c1 <= lim) {
                const uint32_t gw = left ? g0 : g1;
                const uint32_t l2 = gw & 0xFFFFu, r2 = gw >> 16;
                const bool left2 = heap_le(r2, l2);
                H.h[c1] = (uint16_t)(left2 ? l2 : r2);
                i = 2 * c1 + (left2 ? 0 : 1);
                // the next step's words: a step needs i <= lim <= 31 and its
                // second level 2i <= lim, so i <= 15; the clamped reads stay
                // inside the heap's 33 words
                const int g = i < 15 ? i : 15;
                cw = hword(H, i <= lim ? i : 1);
                g0 = hword(H, 2 * g);
                g1 = hword(H, 2 * g + 1);
            }
        }
        if ((len & 1) == 0 && i == len / 2) {  // a last hole with a left child only
            H.h[i] = H.h[2 * i];
            i = 2 * i;
        }
        heap_sift_up1(H, i, value);
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
