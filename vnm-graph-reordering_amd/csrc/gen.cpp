// Seeded synthetic graph generators for the benchmark workloads.
//
// The reference downloads SuiteSparse matrices (scripts/tab5_small.sh:920-941);
// the workloads here are generated instead, from recorded seeds:
//   - Erdos-Renyi: exactly k distinct (r, c) positions drawn by Floyd's
//     sampling over the n*n grid, self loops allowed, values U(lo, hi);
//   - R-MAT: recursive quadrant sampling with (a, b, c, d), self loops
//     rejected, distinct undirected edges until the target count, then
//     mirrored; one U(0,1) value per undirected edge;
//   - Chung-Lu: power-law expected degrees w_i ~ (i + 1)^(-1/(gamma-1)),
//     scaled to the requested mean, endpoints drawn proportionally to w,
//     distinct undirected edges, mirrored.
// Random numbers: splitmix64 seeds xoshiro256**.  Output is a raw tuple list
// in generation order (general storage, both directions for symmetric
// graphs); vnm_coo_from_tuples applies the loader's ordering.
#include "vnm_internal.h"

#include <cmath>
#include <cstring>
#include <unordered_set>
#include <vector>

namespace vnm {

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, 1) with 53 bits
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform float in [0, 1) with 24 bits
    float uniformf() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) {
        // Lemire's nearly divisionless method
        unsigned __int128 m = (unsigned __int128)next() * n;
        uint64_t l = (uint64_t)m;
        if (l < n) {
            uint64_t t = (0 - n) % n;
            while (l < t) {
                m = (unsigned __int128)next() * n;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
};

struct PairHash {
    size_t operator()(uint64_t k) const { return (size_t)(k * 0x9E3779B97F4A7C15ull >> 17); }
};

void emit(std::vector<int32_t> &r, std::vector<int32_t> &c, std::vector<float> &v,
          int32_t a, int32_t b, float x) {
    r.push_back(a);
    c.push_back(b);
    v.push_back(x);
}

}  // namespace

int gen_er(int64_t n, int64_t k, uint64_t seed, float lo, float hi, Tuples &out) {
    if (n <= 0 || k < 0 || k > n * n) return VNM_E_ARG;
    Rng rng(seed);
    const uint64_t total = (uint64_t)n * (uint64_t)n;
    std::unordered_set<uint64_t, PairHash> set;
    set.reserve((size_t)k * 2);
    std::vector<uint64_t> order;
    order.reserve(k);
    for (uint64_t j = total - (uint64_t)k; j < total; ++j) {
        uint64_t t = rng.below(j + 1);
        if (set.insert(t).second) order.push_back(t);
        else { set.insert(j); order.push_back(j); }
    }
    out.nrows = out.ncols = (int32_t)n;
    out.symmetric = false;
    out.row.clear(); out.col.clear(); out.val.clear();
    for (uint64_t p : order)
        emit(out.row, out.col, out.val, (int32_t)(p / n), (int32_t)(p % n),
             lo + (hi - lo) * rng.uniformf());
    return VNM_OK;
}

int gen_rmat(int scale, int64_t undirected_edges, double a, double b, double c,
             uint64_t seed, Tuples &out) {
    if (scale < 1 || scale > 30 || undirected_edges < 0) return VNM_E_ARG;
    const int64_t n = 1ll << scale;
    if (undirected_edges > n * (n - 1) / 2) return VNM_E_ARG;
    Rng rng(seed);
    std::unordered_set<uint64_t, PairHash> set;
    set.reserve((size_t)undirected_edges * 2);
    std::vector<uint64_t> order;
    order.reserve(undirected_edges);
    const double ab = a + b, abc = a + b + c;
    while ((int64_t)order.size() < undirected_edges) {
        uint64_t u = 0, v = 0;
        for (int l = 0; l < scale; ++l) {
            double x = rng.uniform();
            u <<= 1;
            v <<= 1;
            if (x < a) {
            } else if (x < ab) {
                v |= 1;
            } else if (x < abc) {
                u |= 1;
            } else {
                u |= 1;
                v |= 1;
            }
        }
        if (u == v) continue;
        uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
        uint64_t key = (lo << 32) | hi;
        if (set.insert(key).second) order.push_back(key);
    }
    out.nrows = out.ncols = (int32_t)n;
    out.symmetric = false;
    out.row.clear(); out.col.clear(); out.val.clear();
    for (uint64_t key : order) {
        int32_t p = (int32_t)(key >> 32), q = (int32_t)(key & 0xffffffffu);
        float x = rng.uniformf();
        emit(out.row, out.col, out.val, p, q, x);
        emit(out.row, out.col, out.val, q, p, x);
    }
    return VNM_OK;
}

int gen_chung_lu(int64_t n, double gamma, double mean_degree, uint64_t seed, Tuples &out) {
    if (n < 2 || gamma <= 1.0 || mean_degree <= 0) return VNM_E_ARG;
    Rng rng(seed);
    std::vector<double> cdf(n);
    double sum = 0;
    const double ex = -1.0 / (gamma - 1.0);
    for (int64_t i = 0; i < n; ++i) {
        sum += std::pow((double)(i + 1), ex);
        cdf[i] = sum;
    }
    int64_t target = (int64_t)std::llround(mean_degree * (double)n / 2.0);
    if (target > n * (n - 1) / 4) target = n * (n - 1) / 4;
    auto draw = [&]() -> int64_t {
        double x = rng.uniform() * sum;
        int64_t lo = 0, hi = n - 1;
        while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if (cdf[mid] > x) hi = mid; else lo = mid + 1;
        }
        return lo;
    };
    std::unordered_set<uint64_t, PairHash> set;
    set.reserve((size_t)target * 2);
    std::vector<uint64_t> order;
    order.reserve(target);
    int64_t attempts = 0, limit = target * 200 + 1000;
    while ((int64_t)order.size() < target && attempts++ < limit) {
        uint64_t u = (uint64_t)draw(), v = (uint64_t)draw();
        if (u == v) continue;
        uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
        uint64_t key = (lo << 32) | hi;
        if (set.insert(key).second) order.push_back(key);
    }
    out.nrows = out.ncols = (int32_t)n;
    out.symmetric = false;
    out.row.clear(); out.col.clear(); out.val.clear();
    for (uint64_t key : order) {
        int32_t p = (int32_t)(key >> 32), q = (int32_t)(key & 0xffffffffu);
        float x = rng.uniformf();
        emit(out.row, out.col, out.val, p, q, x);
        emit(out.row, out.col, out.val, q, p, x);
    }
    return VNM_OK;
}

void gen_normal(uint64_t seed, int64_t count, float *out) {
    Rng rng(seed);
    for (int64_t i = 0; i < count; i += 2) {
        // Box-Muller
        double u1 = rng.uniform(), u2 = rng.uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = (float)(r * std::cos(6.283185307179586 * u2));
        if (i + 1 < count) out[i + 1] = (float)(r * std::sin(6.283185307179586 * u2));
    }
}

}  // namespace vnm
