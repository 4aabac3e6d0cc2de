"""GPU parity of the allpairs schedules 4..7 (M = 4) against the oracle.

clean mode: stats and relabelled COO bit-exact with the oracle's clean run.
compat mode (the reference's filter and pairtab): bit-exact with the oracle
on every case, including runs where a node is swapped twice in a round: the
reference's relabel is then not a permutation, positions merge, and the run
goes on from the literally relabelled COO (reorder/reorder.h:1194-1197).
"""
import numpy as np
import pytest

import oracle
import vnm

pytestmark = pytest.mark.gpu
N = 2


@pytest.fixture(scope="module")
def ctx():
    c = vnm.Context(0)
    yield c
    c.close()


def _om(m):
    return oracle.Matrix(m.nrows, m.ncols, m.row.copy(), m.col.copy(), m.val.copy())


def _cases():
    rng = np.random.default_rng(44)
    out = []
    for R, d in ((32, 0.3), (64, 0.15), (96, 0.1)):
        mask = rng.random((R, R)) < d
        r, c = np.nonzero(mask)
        out.append(vnm.coo_from_tuples(R, R, r, c, rng.uniform(0.1, 1, r.size).astype(np.float32), 4))
    out.append(vnm.gen_er(1024, 10000, 3, M=4))
    out.append(vnm.gen_rmat(10, 4000, 5, M=4))
    out.append(vnm.gen_rmat(11, 8000, 5, M=4))
    return out


@pytest.mark.parametrize("sched", [4, 5, 6, 7])
def test_allpairs_clean_matches_oracle(ctx, sched):
    for m in _cases():
        om = _om(m)
        ost, ow = oracle.allpairs_reorder(om, N, sched, clean=True)
        assert ow == 0
        g = vnm.Graph(ctx, m, 4)
        st = g.nm_reorder_allpairs(N, sched, clean=True)
        out = g.download()
        assert st == list(ost), (sched, st, list(ost))
        assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col)


@pytest.mark.parametrize("sched", [4, 5, 6, 7])
def test_allpairs_compat_matches_oracle(ctx, sched):
    seen_literal = 0
    for i, m in enumerate(_cases()):
        om = _om(m)
        ost, ow = oracle.allpairs_reorder(om, N, sched, clean=False)
        g = vnm.Graph(ctx, m, 4)
        # sched 4..7 through the reference's entry (the reference filter), and
        # the allpairs entry in compat mode on a second copy
        st = g.nm_reorder(N, sched)
        out = g.download()
        assert st == list(ost), (sched, i, st, list(ost), ow)
        assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col), (sched, i, ow)
        g2 = vnm.Graph(ctx, m, 4)
        st2 = g2.nm_reorder_allpairs(N, sched, clean=False)
        out2 = g2.download()
        assert st2 == st and np.array_equal(out2.row, out.row) and np.array_equal(out2.col, out.col)
        seen_literal += ow > 0
    # the cases reach the literal relabel (a node swapped twice) for every
    # schedule
    assert seen_literal > 0, sched


def test_allpairs_rejects_m8(ctx):
    g = vnm.Graph(ctx, vnm.gen_er(256, 2000, 1), 8)
    with pytest.raises(vnm.VnmError):
        g.nm_reorder_allpairs(N, 4, clean=True)
