"""GPU parity of the allpairs schedules 4..7 (M = 4) against the oracle.

clean mode: stats and relabelled COO bit-exact with the oracle's clean run.
compat mode (the reference's filter and pairtab): bit-exact with the oracle
wherever the oracle reports that no node was swapped twice; where it does,
the reference's relabel is not a permutation and the GPU must refuse with
VNM_E_UB_PATH.
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
def test_allpairs_compat_matches_or_refuses(ctx, sched):
    seen_ok = seen_ub = 0
    for m in _cases():
        om = _om(m)
        ost, ow = oracle.allpairs_reorder(om, N, sched, clean=False)
        g = vnm.Graph(ctx, m, 4)
        if ow == 0:
            st = g.nm_reorder(N, sched)  # sched 4..7 run the reference filter
            out = g.download()
            assert st == list(ost)
            assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col)
            seen_ok += 1
        else:
            with pytest.raises(vnm.VnmError) as e:
                g.nm_reorder(N, sched)
            assert e.value.code == vnm.E_UB_PATH
            seen_ub += 1
    assert seen_ok > 0


def test_allpairs_rejects_m8(ctx):
    g = vnm.Graph(ctx, vnm.gen_er(256, 2000, 1), 8)
    with pytest.raises(vnm.VnmError):
        g.nm_reorder_allpairs(N, 4, clean=True)
