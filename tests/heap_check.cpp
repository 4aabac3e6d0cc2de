// Host check of the device heap emulation (csrc/hip_common.h: heap_push,
// heap_pop over 33-word heaps with sibling pairs in one word) against
// libstdc++'s std::priority_queue with the reference Comparator
// (reorder/swappair.h:112-120, a non-strict less-or-equal on the key; here
// on the dense rank packed in the entry's high byte).  Every heap of 1..64
// entries (M = 1..8), pushed in element order with ranks drawn from 1..MM
// distinct values (ties whenever fewer), then popped empty; the words of the
// neighbouring heaps must stay untouched.  Prints "cases N bad B".
#include "hip_common.h"

#include <cstdio>
#include <queue>
#include <random>
#include <vector>

using namespace vnm;

struct RefLess {
    bool operator()(uint32_t a, uint32_t b) const { return (a >> 8) <= (b >> 8); }
};

int main(int argc, char **argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 300000;
    std::mt19937 rng(7);
    long bad = 0, cases = 0;
    for (long it = 0; it < n; ++it) {
        const int M = 1 + (int)(it % 8), MM = M * M;
        const int nr = 1 + (int)(rng() % MM);
        // three heaps' words: the one under test in the middle
        std::vector<uint32_t> mem(3 * HEAP_WORDS, 0xDEADBEEFu);
        HeapRef H{reinterpret_cast<uint16_t *>(mem.data() + HEAP_WORDS), 0};
        std::priority_queue<uint32_t, std::vector<uint32_t>, RefLess> pq;
        std::vector<uint32_t> vals(MM);
        for (int e = 0; e < MM; ++e) {
            vals[e] = ((uint32_t)(rng() % nr) << 8) | (uint32_t)e;
            hset(H, e, vals[e]);  // staged in slot e, as the greedy kernel does
        }
        for (int e = 0; e < MM; ++e) {
            heap_push(H, hget(H, e));
            pq.push(vals[e]);
        }
        for (int k = 0; k < MM; ++k) {
            const uint32_t a = heap_pop(H), b = pq.top();
            pq.pop();
            if (a != b) {
                ++bad;
                break;
            }
        }
        for (int j = 0; j < HEAP_WORDS; ++j)
            if (mem[j] != 0xDEADBEEFu || mem[2 * HEAP_WORDS + j] != 0xDEADBEEFu) {
                ++bad;
                break;
            }
        ++cases;
    }
    std::printf("cases %ld bad %ld\n", cases, bad);
    return bad != 0;
}
