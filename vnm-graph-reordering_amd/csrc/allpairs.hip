// Allpairs N:M schedules (4..7, M = 4): single and double column swaps.
//
// Replaces run_allpairs / _double / _reverse / _reverse_double
// (reorder/reorder.h:1102-1803) with GetSegSwapScoreAll (:133-260),
// get_segment_swap_score_all (:538-575) and SwapPairs / Comparators
// (reorder/swappair.h:22-98,122-130).  Not reachable from the reference's
// CLI.
//
// Two modes.  compat follows the reference as written: single swaps are
// never dropped by the visited filter (reorder/reorder.h:1194-1197) and a
// double swap applies its pairtab entry, which for moves 33..35 is not the
// move that was scored.  Because singles are never dropped, the reference can
// swap a node that was already swapped in the round; its relabel is then not
// a permutation (two nodes take one id) and the matrix loses positions.
// Such a round is applied literally, as the reference does: the COO is
// brought to its current ids, each id the round's newid table names is
// replaced by it, and the run continues from that COO, its CSC holding every
// position once.  clean drops every move that touches a swapped node and
// applies the move that was scored.
//
// Shape: while no node is swapped twice, each visit of a primary segment
// swaps at least one of its 4 unswapped columns, so a round touches the
// primary and at most the first 4 targets of its traversal.  Per round the
// GPU scores every ordered pair of those <= 5 segments at once: a table pass
// gives each row's bitmaps in the 5 segments, a histogram counts (l, r)
// bitmap pairs per segment pair (256 bins), and each of the 52 moves' gains is
// the histogram weighted by the per-row gain of that move.  A pair outside
// those (a compat round that swaps nodes twice can reach further targets) is
// scored the same way on demand; scores are those of the round's start, as
// the reference's bitmaps change only at its end.  The greedy itself (the
// reference's priority queues, same comparators, same libstdc++ heap) runs on
// the host; swaps are committed to perm/inv and the counts recomputed on the
// device.
#include "graph.h"
#include "hip_common.h"
#include "tables.h"

#include <algorithm>
#include <array>
#include <functional>
#include <map>
#include <queue>
#include <vector>

namespace vnm {

constexpr int AP_K = 5;                   // segments scored per round
constexpr int AP_PAIRS = AP_K * (AP_K - 1) / 2;

// Double swap i: the new l bitmap takes bits AP_POS[i][0..3] and the new r
// bitmap bits AP_POS[i][4..7] of (l0..l3, r0..r3), MSB-first
// (reorder/reorder.h:185-227, kept as data).
__constant__ int8_t c_ap_pos[36][8] = {
    {1, 2, 6, 7, 0, 3, 4, 5}, {1, 2, 5, 7, 0, 3, 4, 6}, {1, 2, 5, 6, 0, 3, 4, 7},
    {1, 2, 4, 7, 0, 3, 5, 6}, {1, 2, 4, 6, 0, 3, 5, 7}, {1, 2, 4, 5, 0, 3, 6, 7},
    {0, 1, 4, 5, 2, 3, 6, 7}, {0, 1, 4, 6, 2, 3, 5, 7}, {0, 1, 4, 7, 2, 3, 5, 6},
    {0, 1, 5, 6, 2, 3, 4, 7}, {0, 1, 5, 7, 2, 3, 4, 6}, {0, 1, 6, 7, 2, 3, 4, 5},
    {0, 2, 4, 5, 1, 3, 6, 7}, {0, 2, 4, 6, 1, 3, 5, 7}, {0, 2, 4, 7, 1, 3, 5, 6},
    {0, 2, 5, 6, 1, 3, 4, 7}, {0, 2, 5, 7, 1, 3, 4, 6}, {0, 2, 6, 7, 1, 3, 4, 5},
    {0, 3, 4, 5, 1, 2, 6, 7}, {0, 3, 4, 6, 1, 2, 5, 7}, {0, 3, 4, 7, 1, 2, 5, 6},
    {0, 3, 5, 6, 1, 2, 4, 7}, {0, 3, 5, 7, 1, 2, 4, 6}, {0, 3, 6, 7, 1, 2, 4, 5},
    {1, 3, 6, 7, 0, 2, 4, 5}, {1, 3, 5, 7, 0, 2, 4, 6}, {1, 3, 5, 6, 0, 2, 4, 7},
    {1, 3, 4, 7, 0, 2, 5, 6}, {1, 3, 4, 6, 0, 2, 5, 7}, {1, 3, 4, 5, 0, 2, 6, 7},
    {2, 3, 4, 5, 0, 1, 6, 7}, {2, 3, 4, 6, 0, 1, 5, 7}, {2, 3, 4, 7, 0, 1, 5, 6},
    {2, 3, 5, 6, 0, 1, 4, 7}, {2, 3, 5, 7, 0, 1, 4, 6}, {2, 3, 6, 7, 0, 1, 4, 5}};

// The reference's applied double swaps (reorder/swappair.h:54-97, data):
// (l a, r b-4) and (l c, r d-4).
static const int h_ap_pairtab[36][4] = {
    {0, 6, 3, 7}, {0, 5, 3, 7}, {0, 5, 3, 6}, {0, 4, 3, 7}, {0, 4, 3, 6}, {0, 4, 3, 5},
    {2, 4, 3, 5}, {2, 4, 3, 6}, {2, 4, 3, 7}, {2, 5, 3, 6}, {2, 5, 3, 7}, {2, 6, 3, 7},
    {1, 4, 3, 5}, {1, 4, 3, 6}, {1, 4, 3, 7}, {1, 5, 3, 6}, {1, 5, 3, 7}, {1, 6, 3, 7},
    {1, 4, 2, 5}, {1, 4, 2, 6}, {1, 4, 2, 7}, {1, 5, 2, 6}, {1, 5, 2, 7}, {1, 6, 2, 7},
    {0, 6, 2, 7}, {0, 5, 2, 7}, {0, 5, 2, 6}, {0, 4, 2, 7}, {0, 4, 2, 6}, {0, 4, 2, 5},
    {0, 4, 1, 5}, {0, 4, 1, 6}, {0, 4, 1, 7}, {0, 5, 1, 7}, {0, 5, 1, 6}, {0, 6, 1, 5}};
static const int h_ap_pos[36][8] = {
    {1, 2, 6, 7, 0, 3, 4, 5}, {1, 2, 5, 7, 0, 3, 4, 6}, {1, 2, 5, 6, 0, 3, 4, 7},
    {1, 2, 4, 7, 0, 3, 5, 6}, {1, 2, 4, 6, 0, 3, 5, 7}, {1, 2, 4, 5, 0, 3, 6, 7},
    {0, 1, 4, 5, 2, 3, 6, 7}, {0, 1, 4, 6, 2, 3, 5, 7}, {0, 1, 4, 7, 2, 3, 5, 6},
    {0, 1, 5, 6, 2, 3, 4, 7}, {0, 1, 5, 7, 2, 3, 4, 6}, {0, 1, 6, 7, 2, 3, 4, 5},
    {0, 2, 4, 5, 1, 3, 6, 7}, {0, 2, 4, 6, 1, 3, 5, 7}, {0, 2, 4, 7, 1, 3, 5, 6},
    {0, 2, 5, 6, 1, 3, 4, 7}, {0, 2, 5, 7, 1, 3, 4, 6}, {0, 2, 6, 7, 1, 3, 4, 5},
    {0, 3, 4, 5, 1, 2, 6, 7}, {0, 3, 4, 6, 1, 2, 5, 7}, {0, 3, 4, 7, 1, 2, 5, 6},
    {0, 3, 5, 6, 1, 2, 4, 7}, {0, 3, 5, 7, 1, 2, 4, 6}, {0, 3, 6, 7, 1, 2, 4, 5},
    {1, 3, 6, 7, 0, 2, 4, 5}, {1, 3, 5, 7, 0, 2, 4, 6}, {1, 3, 5, 6, 0, 2, 4, 7},
    {1, 3, 4, 7, 0, 2, 5, 6}, {1, 3, 4, 6, 0, 2, 5, 7}, {1, 3, 4, 5, 0, 2, 6, 7},
    {2, 3, 4, 5, 0, 1, 6, 7}, {2, 3, 4, 6, 0, 1, 5, 7}, {2, 3, 4, 7, 0, 1, 5, 6},
    {2, 3, 5, 6, 0, 1, 4, 7}, {2, 3, 5, 7, 0, 1, 4, 6}, {2, 3, 6, 7, 0, 1, 4, 5}};

__device__ __forceinline__ int pair_index(int a, int b) {  // a < b < AP_K
    return a * (2 * AP_K - a - 1) / 2 + (b - a - 1);
}

// Fill / clear the K slots of the round's segments (one workgroup each).
__global__ __launch_bounds__(256) void k_ap_slots(int R, const int32_t *__restrict__ segs, int K,
                                                  const int32_t *__restrict__ inv, const int32_t *__restrict__ cptr,
                                                  const int32_t *__restrict__ crow, uint32_t *tab, int clear) {
    __shared__ ColSpan cols[4];
    __shared__ int pre[5];
    __shared__ int sel[1];
    if (blockIdx.x >= K) return;
    if (threadIdx.x == 0) sel[0] = segs[blockIdx.x];
    __syncthreads();
    const int total = table_columns(sel, 1, 4, inv, cptr, cols, pre);
    uint32_t *slot = tab + (size_t)blockIdx.x * (R >> 2);
    if (clear) table_clear(slot, 4, R, cols, pre, 4, total, crow);
    else table_scatter(slot, 4, R, cols, pre, 4, total, crow);
}

// Histogram of (l, r) bitmap pairs for every unordered pair of the K
// segments over the rows present in either.  Segment j's workgroup walks its
// owner entries (one per row) and takes the rows not present in an earlier
// segment, so every row of the union is counted once.
__global__ __launch_bounds__(256) void k_ap_hist(int R, const int32_t *__restrict__ segs, int K,
                                                 const int32_t *__restrict__ inv, const int32_t *__restrict__ cptr,
                                                 const int32_t *__restrict__ crow, uint32_t *tab,
                                                 int32_t *__restrict__ hist) {
    __shared__ ColSpan cols[4];
    __shared__ int pre[5];
    __shared__ int sel[1];
    __shared__ int h[AP_PAIRS * 256];
    const int j = blockIdx.x;
    if (j >= K) return;
    for (int i = threadIdx.x; i < AP_PAIRS * 256; i += 256) h[i] = 0;
    if (threadIdx.x == 0) sel[0] = segs[j];
    __syncthreads();
    const int total = table_columns(sel, 1, 4, inv, cptr, cols, pre);
    for (int e = threadIdx.x; e < total; e += 256) {
        const int t = entry_col(pre, 4, e);
        const int r = crow[cols[t].start + (e - pre[t])];
        uint32_t bm[AP_K];
        bool first = true;
        for (int s = 0; s < K; ++s) {
            bm[s] = slot_byte(tab, s, R, r);
            if (s < j && bm[s]) first = false;
        }
        if (!first || (bm[j] >> (3 - t)) != 1u) continue;  // counted elsewhere / not the owner
        for (int a = 0; a < K; ++a)
            for (int b = a + 1; b < K; ++b)
                if (bm[a] | bm[b]) atomicAdd(&h[pair_index(a, b) * 256 + bm[a] * 16 + bm[b]], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < AP_PAIRS * 256; i += 256)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Per-row gain of move m (0..51) for bitmaps l, r (M = 4): singles m =
// b1*4 + b2 (GetSegSwapScore's per-row rule), doubles m = 16 + i by AP_POS.
__device__ __forceinline__ void ap_row_gain(int m, uint32_t l, uint32_t r, int N, int &lg, int &rg) {
    uint32_t ln, rn;
    if (m < 16) {
        const int b1 = m >> 2, b2 = m & 3;
        const uint32_t x = (l >> (3 - b1)) & 1u, y = (r >> (3 - b2)) & 1u;
        ln = (l & ~(1u << (3 - b1))) | (y << (3 - b1));
        rn = (r & ~(1u << (3 - b2))) | (x << (3 - b2));
    } else {
        const int i = m - 16;
        ln = rn = 0;
        for (int q = 0; q < 4; ++q) {
            const int pl = c_ap_pos[i][q], pr = c_ap_pos[i][4 + q];
            const uint32_t bl = pl < 4 ? (l >> (3 - pl)) & 1u : (r >> (7 - pl)) & 1u;
            const uint32_t br = pr < 4 ? (l >> (3 - pr)) & 1u : (r >> (7 - pr)) & 1u;
            ln |= bl << (3 - q);
            rn |= br << (3 - q);
        }
    }
    lg = (popc32(l) > N) - (popc32(ln) > N);
    rg = (popc32(r) > N) - (popc32(rn) > N);
}

// Gains of every ordered pair (a, b), a != b: score[(a*K + b)][m] = {lgain,
// rgain} of move m for l = bitmap in a, r = bitmap in b.
__global__ void k_ap_gains(int K, int N, const int32_t *__restrict__ hist, int32_t *__restrict__ score) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * K * 52) return;
    const int m = t % 52, ab = t / 52, a = ab / K, b = ab % K;
    if (a == b) return;
    const int lo = min(a, b), hi = max(a, b);
    const int32_t *h = hist + pair_index(lo, hi) * 256;
    int sl = 0, sr = 0;
    for (int x = 0; x < 16; ++x)
        for (int y = 0; y < 16; ++y) {
            const int c = h[x * 16 + y];
            if (!c) continue;
            const uint32_t l = a < b ? x : y, r = a < b ? y : x;
            int lg, rg;
            ap_row_gain(m, l, r, N, lg, rg);
            sl += c * lg;
            sr += c * rg;
        }
    score[(size_t)ab * 104 + 2 * m] = sl;
    score[(size_t)ab * 104 + 2 * m + 1] = sr;
}

// The literal relabel of a compat round: a[i] = newid[a[i]] where that is
// set (reorder/reorder.h:23-35 with the round's newid, which need not be a
// permutation).
__global__ void k_ap_literal(int32_t *__restrict__ a, int64_t n, const int32_t *__restrict__ newid, int32_t R) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = a[i];
        if (x >= 0 && x < R && newid[x] != -1) a[i] = newid[x];
    }
}

// Commit swaps of current ids (pairs (v1, v2)) to perm / inv.
__global__ void k_ap_commit(int n, const int32_t *__restrict__ sw, int32_t *__restrict__ perm, int32_t *__restrict__ inv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int v1 = sw[2 * t], v2 = sw[2 * t + 1];
    const int b1 = inv[v1], b2 = inv[v2];
    inv[v1] = b2;
    inv[v2] = b1;
    perm[b1] = v2;
    perm[b2] = v1;
}

// ------------------------------------------------------------ host greedy

struct ApMove {
    int v1, v2, v3, v4, lg, rg;
};

static bool ap_less(const ApMove *p1, const ApMove *p2) {  // Comparators, reorder/swappair.h:122-130
    if (p1->lg + p1->rg == p2->lg + p2->rg) return p1->rg <= p2->rg;
    return p1->lg + p1->rg < p2->lg + p2->rg;
}

typedef std::pair<int, int> SegCnt;
static bool seg_cmp(const SegCnt &a, const SegCnt &b) {  // cmp, reorder/reorder.h:422-427
    if (a.second == b.second) return a.first > b.first;
    return a.second < b.second;
}
static bool seg_cmp_rev(const SegCnt &a, const SegCnt &b) {  // cmp_reverse, :429-434
    if (a.second == b.second) return a.first < b.first;
    return a.second > b.second;
}
typedef std::priority_queue<SegCnt, std::vector<SegCnt>, std::function<bool(const SegCnt &, const SegCnt &)>> SegList;

int nm_reorder_allpairs(vnm_graph *g, int N, int sched, int maxiter, int clean, vnm_nm_stats *stats) {
    if (g->M != 4) return set_error(VNM_E_ARG, "allpairs schedules need M = 4 (got %d)", g->M);
    if (sched < 4 || sched > 7) return set_error(VNM_E_ARG, "allpairs schedules are 4..7");
    if (N < 1 || N >= 4) return set_error(VNM_E_ARG, "N must be in [1, 4)");
    if (g->nrows % 4 != 0 || g->ncols > g->nrows)
        return set_error(VNM_E_ARG, "N:M reorder needs nrows %% M == 0 and ncols <= nrows");
    const int base = sched - 4;
    const bool rev = base == 2 || base == 3, drain = base == 1 || base == 3;
    int rc;
    if ((rc = build_csc(g))) return rc;
    if ((rc = ensure_tab(g, AP_K))) return rc;
    hipStream_t st = g->ctx->stream;
    const int R = g->nrows, ns = g->nseg();
    DBuf<int32_t> dsegs, dhist, dscore, dsw;
    if ((rc = dsegs.alloc(AP_K)) || (rc = dhist.alloc(AP_PAIRS * 256)) || (rc = dscore.alloc(AP_K * AP_K * 104)) ||
        (rc = dsw.alloc(2 * 64)))
        return rc;
    std::vector<int32_t> cnt(ns), nzr(ns), score(AP_K * AP_K * 104);
    auto recount = [&]() -> int {
        int r = locate_all(g, N, g->cnt.p, g->nzr.p);
        if (r) return r;
        VNM_HIP_TRY(hipMemcpyAsync(cnt.data(), g->cnt.p, ns * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(nzr.data(), g->nzr.p, ns * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        return VNM_OK;
    };
    auto make_list = [&]() {
        SegList l(rev ? seg_cmp_rev : seg_cmp);
        for (int s = 0; s < ns; ++s)
            if (cnt[s] != 0) l.push({s, cnt[s]});
        return l;
    };
    auto total = [](SegList l) {
        int c = 0;
        while (!l.empty()) { c += l.top().second; l.pop(); }
        return c;
    };
    // score every ordered pair of up to AP_K segments; the gains of ordered
    // segment pairs (a, b), 52 moves x {lgain, rgain}, kept for the round
    std::vector<int> ksegs;
    std::map<std::pair<int, int>, std::array<int32_t, 104>> scored;
    auto score_round = [&]() -> int {
        const int K = (int)ksegs.size();
        VNM_HIP_TRY(hipMemcpyAsync(dsegs.p, ksegs.data(), K * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemsetAsync(dhist.p, 0, AP_PAIRS * 256 * 4, st));
        hipLaunchKernelGGL(k_ap_slots, dim3(K), dim3(256), 0, st, R, dsegs.p, K, g->inv.p, g->cptr.p, g->crow.p, g->tab.p, 0);
        hipLaunchKernelGGL(k_ap_hist, dim3(K), dim3(256), 0, st, R, dsegs.p, K, g->inv.p, g->cptr.p, g->crow.p, g->tab.p,
                           dhist.p);
        hipLaunchKernelGGL(k_ap_gains, dim3((K * K * 52 + 255) / 256), dim3(256), 0, st, K, N, dhist.p, dscore.p);
        hipLaunchKernelGGL(k_ap_slots, dim3(K), dim3(256), 0, st, R, dsegs.p, K, g->inv.p, g->cptr.p, g->crow.p, g->tab.p, 1);
        g->ctx->launches += 4;
        VNM_LAUNCH_CHECK("k_ap_gains");
        VNM_HIP_TRY(hipMemcpyAsync(score.data(), dscore.p, (size_t)K * K * 104 * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipStreamSynchronize(st));
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) {
                if (i == j) continue;
                // a segment scored against itself keeps its (0, 1) slot pair
                auto ins = scored.try_emplace({ksegs[i], ksegs[j]});
                if (ins.second) std::copy_n(&score[(size_t)(i * K + j) * 104], 104, ins.first->second.begin());
            }
        return VNM_OK;
    };
    // the moves of segments a x b in push order (get_segment_swap_score_all)
    auto moves = [&](int a, int b, std::vector<ApMove> &out) -> int {
        auto it = scored.find({a, b});
        if (it == scored.end()) {
            ksegs = {a, b};
            int r = score_round();
            if (r) return r;
            it = scored.find({a, b});
        }
        const int32_t *sc = it->second.data();
        out.resize(52);
        for (int i = 0; i < 16; ++i) out[i] = ApMove{a * 4 + i / 4, b * 4 + i % 4, -1, -1, sc[2 * i], sc[2 * i + 1]};
        for (int i = 0; i < 36; ++i) {
            int x, y, u, w;
            if (!clean) {
                x = h_ap_pairtab[i][0]; y = h_ap_pairtab[i][1] - 4; u = h_ap_pairtab[i][2]; w = h_ap_pairtab[i][3] - 4;
            } else {  // the scored exchange, paired in ascending order
                bool kept[4] = {false, false, false, false};
                int taken[2], nt = 0, out_l[2], no = 0;
                for (int q = 0; q < 4; ++q) {
                    if (h_ap_pos[i][q] < 4) kept[h_ap_pos[i][q]] = true;
                    else taken[nt++] = h_ap_pos[i][q] - 4;
                }
                for (int q = 0; q < 4; ++q) if (!kept[q]) out_l[no++] = q;
                std::sort(taken, taken + 2);
                x = out_l[0]; y = taken[0]; u = out_l[1]; w = taken[1];
            }
            out[16 + i] = ApMove{a * 4 + x, b * 4 + y, a * 4 + u, b * 4 + w, sc[2 * (16 + i)], sc[2 * (16 + i) + 1]};
        }
        return VNM_OK;
    };
    std::vector<int> newid(R, -1);
    std::vector<int32_t> swaps;
    bool literal = false;  // this round swapped a node twice: newid is not a permutation
    auto swap_nodes = [&](int a, int b) {
        if (newid[a] != -1 || newid[b] != -1) literal = true;
        newid[a] = b;
        newid[b] = a;
        swaps.push_back(a);
        swaps.push_back(b);
    };
    auto visited = [&](int v) { return v != -1 && newid[v] != -1; };
    auto apply = [&](const ApMove &p) {
        swap_nodes(p.v1, p.v2);
        if (p.v3 != -1 && p.v4 != -1) swap_nodes(p.v3, p.v4);
    };

    if ((rc = recount())) return rc;
    SegList lst = make_list();
    const int init_cnt = total(lst), init_size = (int)lst.size();
    int MAXITER = std::max(maxiter, init_size);
    int round = 1;
    std::vector<ApMove> mv;
    while (round <= MAXITER && lst.size() >= 1) {
        if (lst.size() == 1) {
            // the single invalid segment and the segment with most empty rows
            int top = lst.top().first, minseg = 0;
            for (int s = 1; s < ns; ++s)
                if (R - nzr[s] >= R - nzr[minseg]) minseg = s;
            // when minseg == top the segment is scored against itself, the two
            // bitmaps of a row being independent copies (two slots)
            ksegs = {top, minseg};
            if ((rc = score_round())) return rc;
            if ((rc = moves(top, minseg, mv))) return rc;
            std::priority_queue<const ApMove *, std::vector<const ApMove *>, std::function<bool(const ApMove *, const ApMove *)>>
                pq(ap_less);
            for (auto &m : mv) pq.push(&m);
            apply(*pq.top());  // no visited filter (reorder/reorder.h:1145-1147)
        }
        if (lst.size() >= 2) {
            // the round's segments: the primary and its first 4 targets
            SegList t0 = lst;
            const int primary0 = t0.top().first;
            t0.pop();
            SegList trav(drain ? (rev ? seg_cmp : seg_cmp_rev) : (rev ? seg_cmp_rev : seg_cmp));
            while (!t0.empty()) { trav.push(t0.top()); t0.pop(); }
            ksegs = {primary0};
            while (!trav.empty() && (int)ksegs.size() < AP_K) { ksegs.push_back(trav.top().first); trav.pop(); }
            if ((rc = score_round())) return rc;
        }
        while (lst.size() >= 2) {
            int pseg = lst.top().first, pcnt = lst.top().second;
            lst.pop();
            SegList temp(rev ? seg_cmp_rev : seg_cmp);
            if (drain) {
                SegList t2(rev ? seg_cmp : seg_cmp_rev);
                while (!lst.empty()) { t2.push(lst.top()); lst.pop(); }
                temp = t2;
            } else {
                temp = lst;
            }
            SegList upd(rev ? seg_cmp_rev : seg_cmp);
            while (!temp.empty()) {
                bool all_vis = true;
                for (int i = 0; i < 4; ++i) all_vis = all_vis && visited(pseg * 4 + i);
                if (pcnt <= 0 || all_vis) break;
                const int tseg = temp.top().first;
                int tcnt = temp.top().second;
                if ((rc = moves(pseg, tseg, mv))) return rc;
                std::priority_queue<const ApMove *, std::vector<const ApMove *>, std::function<bool(const ApMove *, const ApMove *)>>
                    pq(ap_less);
                for (auto &m : mv) pq.push(&m);
                if (!clean) {  // QUIRK: singles are never dropped (reorder/reorder.h:1194-1197)
                    while (!pq.empty() && (visited(pq.top()->v1) || visited(pq.top()->v2)) &&
                           (visited(pq.top()->v3) || visited(pq.top()->v4)))
                        pq.pop();
                } else {
                    while (!pq.empty() && (visited(pq.top()->v1) || visited(pq.top()->v2) || visited(pq.top()->v3) ||
                                           visited(pq.top()->v4)))
                        pq.pop();
                }
                if (!pq.empty()) {
                    const ApMove p = *pq.top();
                    apply(p);
                    pcnt -= p.lg;
                    tcnt -= p.rg;
                }
                if (tcnt > 0) upd.push({tseg, tcnt});
                temp.pop();
            }
            lst = upd;
        }
        // reorder(): commit the round's swaps, recount
        const int nsw = (int)swaps.size() / 2;
        if (literal) {
            // the COO at its current ids, then newid applied as it stands
            // (reorder/reorder.h:392-418); the run goes on from that COO
            if ((rc = relabel(g, g->perm.p))) return rc;
            DBuf<int32_t> dnew;
            if ((rc = dnew.alloc(R))) return rc;
            VNM_HIP_TRY(hipMemcpyAsync(dnew.p, newid.data(), (size_t)R * 4, hipMemcpyHostToDevice, st));
            auto grid = [](int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65535))); };
            hipLaunchKernelGGL(k_ap_literal, grid(g->nnz), dim3(256), 0, st, g->row.p, g->nnz, dnew.p, R);
            hipLaunchKernelGGL(k_ap_literal, grid(g->nnz), dim3(256), 0, st, g->col.p, g->nnz, dnew.p, R);
            hipLaunchKernelGGL(k_ap_literal, grid((int64_t)g->perm_total.n), dim3(256), 0, st, g->perm_total.p,
                               (int64_t)g->perm_total.n, dnew.p, R);
            g->ctx->launches += 3;
            VNM_LAUNCH_CHECK("k_ap_literal");
            g->epoch++;
            if ((rc = build_csc_unique(g))) return rc;  // synchronizes: dnew may go
            literal = false;
        } else {
            for (int base_i = 0; base_i < nsw; base_i += 64) {
                const int n = std::min(64, nsw - base_i);
                VNM_HIP_TRY(hipMemcpyAsync(dsw.p, swaps.data() + 2 * base_i, (size_t)n * 8, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(k_ap_commit, dim3(1), dim3(64), 0, st, n, dsw.p, g->perm.p, g->inv.p);
                VNM_HIP_TRY(hipStreamSynchronize(st));
            }
            g->ctx->launches++;
        }
        swaps.clear();
        scored.clear();
        std::fill(newid.begin(), newid.end(), -1);
        if ((rc = recount())) return rc;
        lst = make_list();
        round += 1;
    }
    if ((rc = relabel(g, g->perm.p))) return rc;
    SegList fin = make_list();
    stats->init_size = init_size;
    stats->final_size = (int)fin.size();
    stats->init_cnt = init_cnt;
    stats->final_cnt = total(fin);
    stats->rounds = round - 1;
    return VNM_OK;
}

}  // namespace vnm

using namespace vnm;

extern "C" int vnm_nm_reorder_allpairs(vnm_graph *g, int N, int sched, int maxiter, int clean, vnm_nm_stats *st) {
    if (!g || !st) return set_error(VNM_E_ARG, "bad arguments");
    return nm_reorder_allpairs(g, N, sched, maxiter, clean, st);
}
