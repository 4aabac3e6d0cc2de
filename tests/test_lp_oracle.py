"""Pins of the learned-proposer oracle (oracle/lp_oracle.cpp) on the CPU.

The gradient is checked against central finite differences of the FP64 loss;
the input features against an independent dense count of row popcounts per
segment.  A dropped term, a wrong index or a transposed weight in either
fails these.
"""
import numpy as np
import pytest

import oracle

N, M = 2, 8


def _rand_matrix(R, density, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    A = rng.random((R, R)) < density
    if symmetric:
        A = A | A.T
    r, c = np.nonzero(A)
    val = rng.uniform(0.1, 1.0, r.size).astype(np.float32)
    return oracle.from_tuples(R, R, r.astype(np.int32), c.astype(np.int32), val, M), A


def _dense_features(A, N, M):
    """Independent restatement of the feature definitions from a dense 0/1 matrix."""
    R = A.shape[0]
    nseg = R // M
    pc = A.reshape(R, nseg, M).sum(axis=2)             # pc[r, p]
    cls = np.where(pc < N, 0, np.where(pc == N, 1, np.where(pc == N + 1, 2, 3)))
    S = np.zeros((nseg, 4), np.int64)
    for k in range(4):
        S[:, k] = ((pc > 0) & (cls == k)).sum(axis=0)
    deg = A.sum(axis=0)                                 # rows of column u
    nnzp = deg.reshape(nseg, M).sum(axis=1)
    T = (S[:, 2] + S[:, 3]).sum()
    X = np.zeros((R, 16))
    for u in range(R):
        p = u // M
        rows = np.nonzero(A[:, u])[0]
        ck = np.bincount(cls[rows, p], minlength=4)
        s = S[p]
        dd = max(deg[u], 1)
        X[u] = [1, np.log1p(deg[u]), *(ck / dd), np.log1p(ck[2]), np.log1p(ck[1]), np.log1p(ck[3]),
                np.log1p(s[2] + s[3]), np.log1p(s[1]), np.log1p(s[2]), np.log1p(nnzp[p]),
                float(s[2] + s[3] > 0), np.log1p(s[1] - ck[1]), np.log1p(T) / 16]
    return X


@pytest.mark.parametrize("seed,density,sym", [(1, 0.15, True), (2, 0.3, False), (3, 0.05, True)])
def test_features_dense_count(seed, density, sym):
    m, A = _rand_matrix(48, density, seed, sym)
    X = oracle.lp_features(m, N, M)
    np.testing.assert_allclose(X, _dense_features(A, N, M), rtol=0, atol=1e-12)


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, scale, oracle.LP_NPARAMS)
    P[22720 + 34:] = 0.0          # unused head biases
    for k in range(64):           # unused head columns
        P[18624 + k * 64 + 34:18624 + k * 64 + 64] = 0.0
    return P


def test_gradient_finite_differences():
    m, _ = _rand_matrix(32, 0.2, 7)
    rng = np.random.default_rng(11)
    P = _params(5)
    pairs = rng.integers(0, 32, (24, 2)).astype(np.int32)
    target = rng.normal(0, 1, 24)
    loss, grad, _ = oracle.lp_loss_grad(m, N, M, P, pairs, target)
    assert loss > 0
    # every parameter group: W0 self/neighbour rows, b0, W1, b1, W2, b2, Wh, bh
    groups = [(0, 1024), (1024, 2048), (2048, 2112), (2112, 10304), (10304, 10368),
              (10368, 18560), (18560, 18624), (18624, 22720), (22720, 22754)]
    idx = np.concatenate([rng.integers(a, b, 24) for a, b in groups])
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        lp, _, _ = oracle.lp_loss_grad(m, N, M, Pp, pairs, target, want_grad=False)
        lm, _, _ = oracle.lp_loss_grad(m, N, M, Pm, pairs, target, want_grad=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])
    # unused head outputs carry no gradient
    assert np.all(grad[22720 + 34:] == 0)


def test_pair_score_definition():
    m, _ = _rand_matrix(24, 0.25, 9)
    P = _params(3)
    pairs = np.array([[0, 5], [7, 3], [23, 8]], np.int32)
    _, _, Y = oracle.lp_loss_grad(m, N, M, P, pairs, np.zeros(3), want_grad=False)
    # loss with target = s is zero; with target s + 1 it is one
    s = np.array([Y[u, 0] + Y[v, 1] + Y[u, 2:18] @ Y[v, 18:34] for u, v in pairs])
    l0, _, _ = oracle.lp_loss_grad(m, N, M, P, pairs, s, want_grad=False)
    l1, _, _ = oracle.lp_loss_grad(m, N, M, P, pairs, s + 1, want_grad=False)
    assert abs(l0) < 1e-20 and abs(l1 - 1) < 1e-12
