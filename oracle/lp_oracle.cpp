/*
 * CPU oracle of the learned pair proposer -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 restatement of the proposer network the HIP library trains and runs:
 * the per-node input features, the three mean-aggregation layers, the pair
 * head, the squared-error loss against exact swap gains, and the analytic
 * gradient of that loss with respect to every parameter.  The network is an
 * addition of this project (the reference has no learned component); the
 * integer swap arithmetic it is trained against is the reference's
 * (GetSegSwapScore, reorder/reorder.h:75-131).
 *
 * Pinning (tests/test_lp_oracle.py): the gradient against central finite
 * differences of the loss in FP64; the features against an independent dense
 * numpy count of row popcounts per segment on small matrices.
 *
 * Definitions (shared with the HIP kernels):
 *   graph     node u's list = rows r with a nonzero (r, u) (sorted), deg(u) =
 *             its length; Agg(H)[u] = (1 / max(deg u, 1)) * sum_{r in list} H[r]
 *   classes   for node u in segment p = u / M and a row r of its list, pc =
 *             number of nonzero (r, c) with c in segment p; class 0: pc < N,
 *             1: pc = N, 2: pc = N+1, 3: pc > N+1.  c_k(u) counts u's rows
 *             per class, S_k(p) counts the rows of segment p per class,
 *             nnz_p = sum of deg over p's columns, T = sum_p S_2 + S_3.
 *   features  x0 = 1, x1 = ln(1+deg), x2..5 = c_k / max(deg,1),
 *             x6 = ln(1+c2), x7 = ln(1+c1), x8 = ln(1+c3), x9 = ln(1+S2+S3),
 *             x10 = ln(1+S1), x11 = ln(1+S2), x12 = ln(1+nnz_p),
 *             x13 = [S2+S3 > 0], x14 = ln(1+S1-c1), x15 = ln(1+T)/16
 *   layers    H0 = X; H_{l+1} = relu([H_l, Agg H_l] W_l + b_l), l = 0..2
 *   head      Y = H3 Wh + bh (64 outputs; 0 donor, 1 acceptor, 2..17 P,
 *             18..33 Q, the rest unused)
 *   pair      s(u, v) = Y[u][0] + Y[v][1] + sum_j Y[u][2+j] * Y[v][18+j]
 *   loss      (1/n) sum_i (s(u_i, v_i) - t_i)^2
 * Parameter layout (row-major [k][n]): W0 [32][64] @0, b0 @2048,
 * W1 [128][64] @2112, b1 @10304, W2 [128][64] @10368, b2 @18560,
 * Wh [64][64] @18624, bh @22720; 22784 in total.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int F = 16, H = 64, HD = 64, PQ = 16;
constexpr int OFF_W[3] = {0, 2112, 10368};
constexpr int OFF_B[3] = {2048, 10304, 18560};
constexpr int FIN[3] = {16, 64, 64};
constexpr int OFF_WH = 18624, OFF_BH = 22720;

struct Adj {
    int R = 0;
    std::vector<int> ptr, idx;    // list(u): rows r of nonzero (r, u), sorted
    std::vector<int> tptr, tidx;  // transposed: for row r, the u with r in list(u)
};

void build_adj(const orc_coo *c, Adj &a) {
    const int R = c->nrows;
    a.R = R;
    std::vector<std::pair<int, int>> e;
    e.reserve((size_t)c->nnz);
    for (int64_t i = 0; i < c->nnz; ++i) {
        int r = c->row[i], u = c->col[i];
        if (c->val[i] != 0.0f && r >= 0 && r < R && u >= 0 && u < R) e.push_back({u, r});
    }
    std::sort(e.begin(), e.end());
    a.ptr.assign(R + 1, 0);
    a.idx.resize(e.size());
    for (size_t i = 0; i < e.size(); ++i) { a.ptr[e[i].first + 1]++; a.idx[i] = e[i].second; }
    for (int u = 0; u < R; ++u) a.ptr[u + 1] += a.ptr[u];
    std::vector<std::pair<int, int>> t(e.size());
    for (size_t i = 0; i < e.size(); ++i) t[i] = {e[i].second, e[i].first};
    std::sort(t.begin(), t.end());
    a.tptr.assign(R + 1, 0);
    a.tidx.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i) { a.tptr[t[i].first + 1]++; a.tidx[i] = t[i].second; }
    for (int r = 0; r < R; ++r) a.tptr[r + 1] += a.tptr[r];
}

void features(const orc_coo *c, const Adj &a, int N, int M, std::vector<double> &X) {
    const int R = a.R, nseg = (R + M - 1) / M;
    // pc of (row, segment) from the row-major lists
    std::vector<std::vector<std::pair<int, int>>> rowseg(R);  // (seg, pc)
    for (int r = 0; r < R; ++r) {
        std::vector<int> segs;
        for (int k = a.tptr[r]; k < a.tptr[r + 1]; ++k) segs.push_back(a.tidx[k] / M);
        std::sort(segs.begin(), segs.end());
        for (size_t i = 0; i < segs.size();) {
            size_t j = i;
            while (j < segs.size() && segs[j] == segs[i]) ++j;
            rowseg[r].push_back({segs[i], (int)(j - i)});
            i = j;
        }
    }
    auto pc_of = [&](int r, int p) {
        for (auto &q : rowseg[r]) if (q.first == p) return q.second;
        return 0;
    };
    auto cls = [&](int pc) { return pc < N ? 0 : pc == N ? 1 : pc == N + 1 ? 2 : 3; };
    std::vector<long> S((size_t)nseg * 4, 0), nnzp(nseg, 0);
    for (int r = 0; r < R; ++r)
        for (auto &q : rowseg[r]) S[(size_t)q.first * 4 + cls(q.second)]++;
    for (int u = 0; u < R; ++u) nnzp[u / M] += a.ptr[u + 1] - a.ptr[u];
    long T = 0;
    for (int p = 0; p < nseg; ++p) T += S[p * 4 + 2] + S[p * 4 + 3];
    X.assign((size_t)R * F, 0.0);
    for (int u = 0; u < R; ++u) {
        const int p = u / M;
        long ck[4] = {0, 0, 0, 0};
        const int deg = a.ptr[u + 1] - a.ptr[u];
        for (int k = a.ptr[u]; k < a.ptr[u + 1]; ++k) ck[cls(pc_of(a.idx[k], p))]++;
        const long *s = &S[(size_t)p * 4];
        double *x = &X[(size_t)u * F];
        const double dd = std::max(deg, 1);
        x[0] = 1.0;
        x[1] = std::log1p((double)deg);
        for (int k = 0; k < 4; ++k) x[2 + k] = ck[k] / dd;
        x[6] = std::log1p((double)ck[2]);
        x[7] = std::log1p((double)ck[1]);
        x[8] = std::log1p((double)ck[3]);
        x[9] = std::log1p((double)(s[2] + s[3]));
        x[10] = std::log1p((double)s[1]);
        x[11] = std::log1p((double)s[2]);
        x[12] = std::log1p((double)nnzp[p]);
        x[13] = (s[2] + s[3]) > 0 ? 1.0 : 0.0;
        x[14] = std::log1p((double)(s[1] - ck[1]));
        x[15] = std::log1p((double)T) / 16.0;
    }
}

// out[u][:] (+)= scale(u) * sum_{j in list(u)} pre(j) * Hin[j][:]
void aggregate(const std::vector<int> &ptr, const std::vector<int> &idx, const double *Hin, int W,
               const std::vector<double> *pre, const std::vector<double> *post, double *out,
               bool acc) {
    const int R = (int)ptr.size() - 1;
    for (int u = 0; u < R; ++u) {
        std::vector<double> s(W, 0.0);
        for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
            int j = idx[k];
            double f = pre ? (*pre)[j] : 1.0;
            for (int w = 0; w < W; ++w) s[w] += f * Hin[(size_t)j * W + w];
        }
        double g = post ? (*post)[u] : 1.0;
        for (int w = 0; w < W; ++w) {
            double v = g * s[w];
            out[(size_t)u * W + w] = acc ? out[(size_t)u * W + w] + v : v;
        }
    }
}

}  // namespace

extern "C" int orc_lp_features(const orc_coo *c, int N, int M, double *X) {
    if (!c || !X || M < 2 || M > 8 || N < 1 || N >= M || c->nrows % M) return ORC_E_ARG;
    Adj a;
    build_adj(c, a);
    std::vector<double> x;
    features(c, a, N, M, x);
    std::copy(x.begin(), x.end(), X);
    return ORC_OK;
}

extern "C" int orc_lp_loss_grad(const orc_coo *c, int N, int M, const double *P,
                                const int32_t *pairs, const double *target, int count,
                                double *loss, double *grad, double *head) {
    if (!c || !P || !loss || count < 0 || (count > 0 && (!pairs || !target)) || M < 2 || M > 8 ||
        N < 1 || N >= M || c->nrows % M)
        return ORC_E_ARG;
    Adj a;
    build_adj(c, a);
    const int R = a.R;
    for (int i = 0; i < 2 * count; ++i)
        if (pairs[i] < 0 || pairs[i] >= R) return ORC_E_ARG;
    std::vector<double> inv_deg(R);
    for (int u = 0; u < R; ++u) inv_deg[u] = 1.0 / std::max(a.ptr[u + 1] - a.ptr[u], 1);

    // forward
    std::vector<double> Hs[4], As[3];
    features(c, a, N, M, Hs[0]);
    for (int l = 0; l < 3; ++l) {
        const int fi = FIN[l];
        As[l].assign((size_t)R * fi, 0.0);
        aggregate(a.ptr, a.idx, Hs[l].data(), fi, nullptr, &inv_deg, As[l].data(), false);
        Hs[l + 1].assign((size_t)R * H, 0.0);
        const double *W = P + OFF_W[l], *b = P + OFF_B[l];
        for (int u = 0; u < R; ++u)
            for (int n = 0; n < H; ++n) {
                double s = b[n];
                for (int k = 0; k < fi; ++k) s += Hs[l][(size_t)u * fi + k] * W[k * H + n];
                for (int k = 0; k < fi; ++k) s += As[l][(size_t)u * fi + k] * W[(fi + k) * H + n];
                Hs[l + 1][(size_t)u * H + n] = s > 0 ? s : 0.0;
            }
    }
    std::vector<double> Y((size_t)R * HD);
    for (int u = 0; u < R; ++u)
        for (int n = 0; n < HD; ++n) {
            double s = P[OFF_BH + n];
            for (int k = 0; k < H; ++k) s += Hs[3][(size_t)u * H + k] * P[OFF_WH + k * HD + n];
            Y[(size_t)u * HD + n] = s;
        }
    if (head) std::copy(Y.begin(), Y.end(), head);

    double L = 0.0;
    std::vector<double> dY((size_t)R * HD, 0.0);
    for (int i = 0; i < count; ++i) {
        const int u = pairs[2 * i], v = pairs[2 * i + 1];
        const double *yu = &Y[(size_t)u * HD], *yv = &Y[(size_t)v * HD];
        double s = yu[0] + yv[1];
        for (int j = 0; j < PQ; ++j) s += yu[2 + j] * yv[2 + PQ + j];
        const double d = s - target[i];
        L += d * d;
        const double ds = count ? 2.0 * d / count : 0.0;
        dY[(size_t)u * HD + 0] += ds;
        dY[(size_t)v * HD + 1] += ds;
        for (int j = 0; j < PQ; ++j) {
            dY[(size_t)u * HD + 2 + j] += ds * yv[2 + PQ + j];
            dY[(size_t)v * HD + 2 + PQ + j] += ds * yu[2 + j];
        }
    }
    *loss = count ? L / count : 0.0;
    if (!grad) return ORC_OK;

    // backward
    std::fill(grad, grad + ORC_LP_NPARAMS, 0.0);
    std::vector<double> dH((size_t)R * H, 0.0);
    for (int u = 0; u < R; ++u)
        for (int n = 0; n < HD; ++n) {
            const double g = dY[(size_t)u * HD + n];
            if (g == 0.0) continue;
            grad[OFF_BH + n] += g;
            for (int k = 0; k < H; ++k) {
                grad[OFF_WH + k * HD + n] += Hs[3][(size_t)u * H + k] * g;
                dH[(size_t)u * H + k] += g * P[OFF_WH + k * HD + n];
            }
        }
    for (int l = 2; l >= 0; --l) {
        const int fi = FIN[l];
        const double *W = P + OFF_W[l];
        std::vector<double> dPre((size_t)R * H);
        for (size_t i = 0; i < dPre.size(); ++i) dPre[i] = Hs[l + 1][i] > 0 ? dH[i] : 0.0;
        std::vector<double> dIn((size_t)R * fi, 0.0), dAg((size_t)R * fi, 0.0);
        for (int u = 0; u < R; ++u)
            for (int n = 0; n < H; ++n) {
                const double g = dPre[(size_t)u * H + n];
                if (g == 0.0) continue;
                grad[OFF_B[l] + n] += g;
                for (int k = 0; k < fi; ++k) {
                    grad[OFF_W[l] + k * H + n] += Hs[l][(size_t)u * fi + k] * g;
                    grad[OFF_W[l] + (fi + k) * H + n] += As[l][(size_t)u * fi + k] * g;
                    dIn[(size_t)u * fi + k] += g * W[k * H + n];
                    dAg[(size_t)u * fi + k] += g * W[(fi + k) * H + n];
                }
            }
        if (l == 0) break;
        // dH_l = dIn + Agg^T dAg, Agg^T(G)[r] = sum_{u : r in list(u)} G[u] / max(deg u, 1)
        aggregate(a.tptr, a.tidx, dAg.data(), fi, &inv_deg, nullptr, dIn.data(), true);
        dH.swap(dIn);
    }
    return ORC_OK;
}
