"""GPU parity of the N:M engine against the oracle (bit-exact integers).

Covers the locate kernel, the batched ballot/popcount scorer, the device
libstdc++ heap emulation and the persistent greedy for schedules 0-3 and the
plain `run` schedule, on random small matrices, the reference fixtures and
seeded ER / R-MAT graphs.
"""
import json
import os

import numpy as np
import pytest

import oracle
import vnm

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = vnm.Context(0)
    yield c
    c.close()


def _omat(m):
    return oracle.Matrix(m.nrows, m.ncols, m.row, m.col, m.val)


def _rand(rng, R, dens, M, sym=False, zero_frac=0.05):
    mask = rng.random((R, R)) < dens
    if sym:
        mask |= mask.T
    r, c = np.nonzero(mask)
    v = rng.uniform(-1, 1, r.shape[0]).astype(np.float32)
    v[rng.random(r.shape[0]) < zero_frac] = 0.0
    return vnm.coo_from_tuples(R, R, r, c, v, M)


def test_heap_pop_order_matches_reference_comparator(ctx):
    recs = json.load(open(os.path.join(GOLD, "ref_pq.json")))
    for M in (4, 8):
        rs = [r for r in recs if r["M"] == M]
        got = ctx.pq_pop_order(np.array([r["score"] for r in rs]), M)
        for g, r in zip(got, rs):
            assert g.tolist() == r["order"]


def test_locate_matches_oracle(ctx):
    rng = np.random.default_rng(1)
    cases = [(_rand(rng, int(rng.integers(1, 20)) * M, rng.uniform(0.02, 0.5), M), M)
             for M in (4, 8) for _ in range(6)]
    cases.append((vnm.gen_er(2048, 41943, 1), 8))
    cases.append((vnm.gen_rmat(13, 40000, 5), 8))
    cases.append((vnm.gen_rmat(12, 30000, 6, M=4), 4))
    for m, M in cases:
        g = vnm.Graph(ctx, m, M)
        for N in (1, 2, 3):
            if N >= M:
                continue
            cnt, zero = g.locate(N)
            ocnt, ozero = oracle.locate(_omat(m), N, M)
            assert np.array_equal(cnt, ocnt) and np.array_equal(zero, ozero)


def test_swap_score_matches_oracle(ctx):
    rng = np.random.default_rng(2)
    cases = [(_rand(rng, int(rng.integers(2, 20)) * M, rng.uniform(0.05, 0.6), M), M)
             for M in (4, 8) for _ in range(4)]
    cases.append((vnm.gen_rmat(12, 30000, 7), 8))
    cases.append((vnm.gen_er(2048, 41943, 1), 8))
    for m, M in cases:
        g = vnm.Graph(ctx, m, M)
        ns = g.nseg
        om = _omat(m)
        for N in (1, 2):
            seg1 = int(rng.integers(0, ns))
            seg2 = rng.integers(0, ns, size=min(ns, 70))
            got = g.swap_score(N, seg1, seg2)
            for k, s2 in enumerate(seg2.tolist()):
                assert np.array_equal(got[k], oracle.swap_score(om, N, M, seg1, s2)), (seg1, s2)


def _nm_case(ctx, m, M, N, sched, maxiter=10):
    g = vnm.Graph(ctx, m, M)
    st = g.nm_reorder(N, sched, maxiter)
    out = g.download()
    om = _omat(m).copy()
    ost, _ = oracle.nm_reorder(om, N, M, sched, maxiter)
    assert st == ost.tolist(), (sched, st, ost.tolist())
    assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col)
    assert np.array_equal(out.val, om.val)
    return st


@pytest.mark.parametrize("sched", [0, 1, 2, 3, 8])
def test_greedy_random_small(ctx, sched):
    rng = np.random.default_rng(10 + sched)
    for t in range(12):
        M = (4, 8)[t % 2]
        N = 2 if M == 8 else int(rng.integers(1, 3))
        R = int(rng.integers(2, 30)) * M
        m = _rand(rng, R, rng.uniform(0.03, 0.5), M, sym=t % 3 == 0)
        _nm_case(ctx, m, M, N, sched)


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_greedy_fixtures(ctx, sched):
    for f in ("GD01_c", "GD95_a", "lund_b", "bcsstm02", "small"):
        m = vnm.mtx_read(os.path.join(GOLD, "mtx", f + ".mtx"), 8)
        if m.nnz:
            _nm_case(ctx, m, 8, 2, sched)


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_greedy_er_2048(ctx, sched):
    # the 2,048-node ER workload (41,943 entries, U(-1,1), seed 1)
    st = _nm_case(ctx, vnm.gen_er(2048, 41943, 1), 8, 2, sched)
    assert st[0] == 22


def test_greedy_rmat_medium(ctx):
    # R-MAT scale 12: hubs make the greedy run many rounds
    m = vnm.gen_rmat(12, 32768, 11)
    for sched in (0, 3):
        _nm_case(ctx, m, 8, 2, sched)
    _nm_case(ctx, vnm.gen_rmat(11, 16384, 12, M=4), 4, 2, 0)


def _batch_cases(rng):
    cases = []
    for t in range(10):
        M = (4, 8)[t % 2]
        R = int(rng.integers(2, 40)) * M
        cases.append((_rand(rng, R, rng.uniform(0.03, 0.5), M, sym=t % 3 == 0), M))
    for f in ("GD01_c", "lund_b", "temp"):  # temp.mtx is already valid: init_size 0
        cases.append((vnm.mtx_read(os.path.join(GOLD, "mtx", f + ".mtx"), 8), 8))
    cases.append((vnm.coo_from_tuples(16, 16, [], [], [], 8), 8))  # no entries
    cases.append((vnm.gen_er(2048, 41943, 1), 8))
    cases.append((vnm.gen_rmat(12, 32768, 11), 8))
    return cases


def test_greedy_batch_matches_oracle(ctx):
    # uneven graphs, mixed M and schedules in one launch chain; each result
    # must be the oracle's for the graph alone, and a second batch and a
    # single run on the same graphs continue from their reordered state
    rng = np.random.default_rng(40)
    cases = _batch_cases(rng)
    scheds = [(0, 1, 2, 3, 8)[i % 5] for i in range(len(cases))]
    graphs = [vnm.Graph(ctx, m, M) for m, M in cases]
    oms = [_omat(m).copy() for m, _ in cases]
    for rep in range(2):
        got = vnm.Graph.nm_reorder_batch(graphs, 2, scheds)
        for i, ((m, M), g, om) in enumerate(zip(cases, graphs, oms)):
            ost, _ = oracle.nm_reorder(om, 2, M, scheds[i], 10)
            assert got[i] == ost.tolist(), (rep, i, scheds[i], got[i], ost.tolist())
            out = g.download()
            assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col), (rep, i)
    m, M = cases[-1]
    st = graphs[-1].nm_reorder(2, 0)
    ost, _ = oracle.nm_reorder(oms[-1], 2, M, 0, 10)
    assert st == ost.tolist()
    out = graphs[-1].download()
    assert np.array_equal(out.row, oms[-1].row) and np.array_equal(out.col, oms[-1].col)


def test_greedy_batch_rejects_bad_batches(ctx):
    m = vnm.gen_er(256, 2000, 3)
    a, b = vnm.Graph(ctx, m, 8), vnm.Graph(ctx, m, 8)
    with pytest.raises(vnm.VnmError):
        vnm.Graph.nm_reorder_batch([a, a], 2, 0)  # the same graph twice
    with pytest.raises(vnm.VnmError):
        vnm.Graph.nm_reorder_batch([a, b], 2, [0, 5])  # allpairs schedules are not batched
    with pytest.raises(vnm.VnmError):
        vnm.Graph.nm_reorder_batch([a, b], 8, 0)  # N >= M
    other = vnm.Context(0)
    try:
        c = vnm.Graph(other, m, 8)
        with pytest.raises(vnm.VnmError):
            vnm.Graph.nm_reorder_batch([a, c], 2, 0)  # graphs of two contexts
        c.close()
    finally:
        other.close()
    # the rejected calls left the graphs usable
    got = vnm.Graph.nm_reorder_batch([a, b], 2, [0, 3])
    want = [oracle.nm_reorder(_omat(m).copy(), 2, 8, s, 10)[0].tolist() for s in (0, 3)]
    assert got == want


def test_greedy_batch_fullsize_golden(ctx):
    # find_best_2toX's four schedules on R-MAT 65,536 (seed 2) in one batch:
    # each graph's stats and relabelled COO equal the oracle's own run
    import hashlib
    gold = json.load(open(os.path.join(GOLD, "fullsize.json")))
    m = vnm.gen_rmat(16, 524288, 2)
    base = vnm.Graph(ctx, m, 8)
    gs = [base.clone() for _ in range(4)]
    got = vnm.Graph.nm_reorder_batch(gs, 2, [0, 1, 2, 3], 10)
    for s, (g, st) in enumerate(zip(gs, got)):
        ref = gold["rmat16_seed2_nm_sched%d" % s]
        assert st == ref["stats"], (s, st)
        out = g.download()
        h = hashlib.sha256()
        for a in (out.row, out.col, out.val):
            h.update(np.ascontiguousarray(a).tobytes())
        assert h.hexdigest() == ref["sha256"], s
    imp = [(r[2] - r[3]) / r[2] for r in (gold["rmat16_seed2_nm_sched%d" % s]["stats"] for s in range(4))]
    assert base.find_best_sched(2, 4) == int(np.argmax(imp))


def test_kat_temp_mtx(ctx):
    m = vnm.mtx_read(os.path.join(GOLD, "mtx", "temp.mtx"), 8)
    g = vnm.Graph(ctx, m, 8)
    assert g.nm_reorder(2, 0) == [0, 0, 0, 0, 0]
    out = g.download()
    assert np.array_equal(out.row, m.row) and np.array_equal(out.col, m.col)


def test_greedy_rmat_fullsize_golden(ctx):
    # R-MAT 65,536 nodes, nnz 1,048,576 (seed 2): stats and relabelled COO
    # equal the oracle's run, recorded by tests/golden/make_fullsize_golden.py
    import hashlib
    gold = json.load(open(os.path.join(GOLD, "fullsize.json")))["rmat16_seed2_nm_sched0"]
    m = vnm.gen_rmat(16, 524288, 2)
    g = vnm.Graph(ctx, m, 8)
    assert g.nm_reorder(2, 0, 10) == gold["stats"]
    out = g.download()
    h = hashlib.sha256()
    for a in (out.row, out.col, out.val):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == gold["sha256"]


def test_context_outlives_a_garbage_cycle_of_its_graphs():
    """Graphs kept only by a reference cycle are finalized by the cyclic
    collector after their context was closed: the context's handle goes with
    the last of them, never before one of them has destroyed its own."""
    import gc
    c = vnm.Context(0)
    m = vnm.gen_er(256, 2000, 5)
    a, b = vnm.Graph(c, m, 8), vnm.Graph(c, m, 8)
    a.other, b.other = b, a  # a cycle: only the cyclic collector frees it
    a.nm_reorder(2, 0)
    c.close()
    assert c.h is not None  # two graphs are still alive
    del a, b
    gc.collect()
    assert c.h is None
