"""CPU tests of the oracle against the reference's own artefacts.

Pins (see oracle/oracle.cpp header): the loader against the reference's
loader compiled from its sources (golden ref_load.json), the SwapPairList pop
order against the reference comparator in std::priority_queue (ref_pq.json),
the VENOM trajectory against reordering.py itself (venom_ref.json), the gain
arithmetic against the closed form, the greedy against a literal dense
restatement, and the temp.mtx / output.txt known answer.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import literal_ref as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MTX = os.path.join(GOLD, "mtx")


def _digest(m):
    h = hashlib.sha256()
    h.update(np.array([m.nrows, m.ncols, m.nnz], np.int64).tobytes())
    h.update(m.row.astype(np.int32).tobytes())
    h.update(m.col.astype(np.int32).tobytes())
    h.update(m.val.astype(np.float32).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def test_loader_matches_reference_loader():
    # the reference's fixtures and the loader corner cases (repeated
    # positions, symmetric with self-loops and both triangles, integer and
    # pattern fields, padding, rectangular): the COO the reference's own
    # loader leaves, compiled from its sources (oracle/refbuild)
    recs = json.load(open(os.path.join(GOLD, "ref_load.json")))
    assert {r["file"] for r in recs if r["dir"] == "mtx_corners"} >= {
        "dup_general.mtx", "dup_many.mtx", "sym_selfloops.mtx", "sym_both_triangles.mtx",
        "integer_general.mtx", "pattern_symmetric.mtx", "pad_rect.mtx"}
    for rec in recs:
        m = oracle.mtx_read(os.path.join(GOLD, rec["dir"], rec["file"]), rec["M"])
        assert (m.nrows, m.ncols, m.nnz) == (rec["nrows"], rec["ncols"], rec["nnz"]), rec["file"]
        assert _digest(m) == rec["sha256"], rec["file"]
        if "row" in rec:
            assert m.row.tolist() == rec["row"]
            assert m.col.tolist() == rec["col"]
            assert np.array_equal(m.val, np.array(rec["val"], np.float32))


def test_pq_order_matches_reference_comparator():
    for rec in json.load(open(os.path.join(GOLD, "ref_pq.json"))):
        got = oracle.pq_pop_order(np.array(rec["score"]), rec["M"]).tolist()
        assert got == rec["order"]
        assert L.pop_order(rec["score"], rec["M"]) == rec["order"]


def test_venom_matches_reference_script():
    for c in json.load(open(os.path.join(GOLD, "venom_ref.json"))):
        m = oracle.Matrix(c["nrows"], c["ncols"], c["row"], c["col"], c["val"])
        assert list(oracle.venom_eval(m, c["M"], c["V"])) == c["evalonly"]
        st = oracle.venom_reorder(m, c["M"], c["V"], 10)
        assert st.tolist() == c["output"], (c["name"], c["V"])
        pos = sorted(set(zip(m.row.tolist(), m.col.tolist())))
        assert [list(p) for p in pos] == c["out_pos"]


def test_temp_mtx_known_answer():
    # reference fixture: output.txt = "0,0" for --evalonly --m 8 --v <= 32
    m = oracle.mtx_read(os.path.join(MTX, "temp.mtx"), 8)
    assert open(os.path.join(MTX, "temp_output.txt")).read().strip() == "0,0"
    for V in (1, 8, 16, 32):
        assert oracle.venom_eval(m, 8, V) == (0, 0)
    cnt, zero = oracle.locate(m, 2, 8)
    assert int((cnt != 0).sum()) == 0
    a = m.copy()
    st, _ = oracle.nm_reorder(a, 2, 8, 0, 10)
    assert st.tolist() == [0, 0, 0, 0, 0]
    assert np.array_equal(a.row, m.row) and np.array_equal(a.col, m.col)


def _closed_form(l, r, b1, b2, N, M):
    """Per-row gain, closed form: only the exchanged bits matter."""
    x = (l >> (M - 1 - b1)) & 1
    y = (r >> (M - 1 - b2)) & 1
    pl, pr = bin(l).count("1"), bin(r).count("1")
    if x == y:
        return 0, 0
    if x == 1:
        return int(pl == N + 1), -int(pr == N)
    return -int(pl == N), int(pr == N + 1)


@pytest.mark.parametrize("M,N", [(4, 2), (8, 2), (4, 1), (8, 3)])
def test_swap_gain_closed_form_exhaustive(M, N):
    # every (l, r) bitmap pair as one row: seg 0 holds l, seg 1 holds r
    pairs = [(l, r) for l in range(1 << M) for r in range(1 << M)]
    rng = np.random.default_rng(M * 10 + N)
    rows, cols = [], []
    for i, (l, r) in enumerate(pairs):
        for j in range(M):
            if (l >> (M - 1 - j)) & 1:
                rows.append(i); cols.append(j)
            if (r >> (M - 1 - j)) & 1:
                rows.append(i); cols.append(M + j)
    R = len(pairs)
    m = oracle.from_tuples(R, R, np.array(rows), np.array(cols), np.ones(len(rows)), M)
    got = oracle.swap_score(m, N, M, 0, 1).reshape(M * M, 2)
    exp = np.zeros((M * M, 2), np.int64)
    for (l, r) in pairs:
        for b1 in range(M):
            for b2 in range(M):
                g = _closed_form(l, r, b1, b2, N, M)
                exp[b1 * M + b2] += g
    assert np.array_equal(got, exp)
    # random row subsets, so a per-row error cannot cancel in the full sum
    for _ in range(5):
        keep = rng.random(R) < 0.3
        idx = np.nonzero(keep)[0]
        sel = np.isin(np.array(rows), idx)
        m2 = oracle.from_tuples(R, R, np.array(rows)[sel], np.array(cols)[sel],
                                np.ones(int(sel.sum())), M)
        got2 = oracle.swap_score(m2, N, M, 0, 1).reshape(M * M, 2)
        exp2 = np.zeros((M * M, 2), np.int64)
        for i in idx:
            l, r = pairs[i]
            for b1 in range(M):
                for b2 in range(M):
                    exp2[b1 * M + b2] += _closed_form(l, r, b1, b2, N, M)
        assert np.array_equal(got2, exp2)


def _rand(rng, R, dens, M, sym=False, zero_frac=0.05):
    mask = rng.random((R, R)) < dens
    if sym:
        mask |= mask.T
    r, c = np.nonzero(mask)
    v = rng.uniform(-1, 1, r.shape[0]).astype(np.float32)
    v[rng.random(r.shape[0]) < zero_frac] = 0.0
    return oracle.from_tuples(R, R, r, c, v, M)


def test_locate_brute_force():
    rng = np.random.default_rng(7)
    for t in range(20):
        M = (4, 8)[t % 2]
        R = int(rng.integers(1, 12)) * M
        m = _rand(rng, R, rng.uniform(0.05, 0.7), M)
        cnt, zero = oracle.locate(m, 2, M)
        bm = L.dense_bitmaps(R, M, m.row.tolist(), m.col.tolist(), m.val.tolist())
        for s in range(R // M):
            assert cnt[s] == sum(1 for row in bm if bin(row[s]).count("1") > 2)
            assert zero[s] == sum(1 for row in bm if row[s] == 0)


@pytest.mark.parametrize("sched", [0, 1, 2, 3, 8])
def test_greedy_matches_literal_restatement(sched):
    rng = np.random.default_rng(100 + sched)
    for t in range(10):
        M = (4, 8)[t % 2]
        N = 2 if M == 8 else int(rng.integers(1, 3))
        R = int(rng.integers(2, 7)) * M
        m = _rand(rng, R, rng.uniform(0.1, 0.6), M, sym=t % 3 == 0)
        a = m.copy()
        st, _ = oracle.nm_reorder(a, N, M, sched, 10)
        rows, cols = [int(x) for x in m.row], [int(x) for x in m.col]
        exp = L.nm_reorder(R, rows, cols, [float(x) for x in m.val], N, M, sched, 10)
        assert st.tolist() == exp
        assert a.row.tolist() == rows and a.col.tolist() == cols
    assert oracle.ub_events() == 0


def test_greedy_invariants():
    rng = np.random.default_rng(5)
    for t in range(12):
        M = (4, 8)[t % 2]
        R = int(rng.integers(4, 30)) * M
        m = _rand(rng, R, rng.uniform(0.02, 0.3), M, sym=t % 2 == 0)
        for sched in (0, 1, 2, 3):
            a = m.copy()
            st, trace = oracle.nm_reorder(a, 2, M, sched, 10, trace_cap=100000)
            # relabel is one symmetric permutation applied to rows and columns
            perm = -np.ones(R, np.int64)
            for src, dst in ((m.row, a.row), (m.col, a.col)):
                for s, d in zip(src.tolist(), dst.tolist()):
                    assert perm[s] in (-1, d)
                    perm[s] = d
            used = perm[perm >= 0]
            assert len(set(used.tolist())) == len(used)
            assert np.array_equal(a.val, m.val)
            # reported final count equals a fresh recount of the output
            cnt, _ = oracle.locate(a, 2, M)
            assert st[1] == int((cnt != 0).sum()) and st[3] == int(cnt.sum())
            cnt0, _ = oracle.locate(m, 2, M)
            assert st[0] == int((cnt0 != 0).sum()) and st[2] == int(cnt0.sum())
            assert trace is not None
    assert oracle.ub_events() == 0


def test_writer_format_and_roundtrip(tmp_path):
    m = oracle.from_tuples(10, 10, np.array([0, 3, 9]), np.array([1, 2, 9]),
                           np.array([0.125, -0.004, 2.5]), 8)
    p = tmp_path / "w.mtx"
    oracle.mtx_write(p, m)
    assert p.read_text() == ("%%MatrixMarket matrix coordinate real general\n"
                             "16 16 3\n1 2 0.12\n4 3 -0.00\n10 10 2.50\n")
    back = oracle.mtx_read(p, 8)
    assert (back.nrows, back.nnz) == (16, 3)
    assert back.val.tolist() == [np.float32(0.12), np.float32(-0.0), np.float32(2.5)]


def test_quantisation_roundtrips():
    for v, a, b in [(0.004, 0.0, 0.0), (0.126, 0.13, 0.13), (1.25, 1.25, 1.2),
                    (12.345, 12.35, 12.0), (-0.994, -0.99, -0.99)]:
        assert oracle.quant_cpp_write(v) == np.float32(a)
        assert oracle.quant_py_roundtrip(v) == np.float32(b)


def test_duplicate_value_quirk():
    # duplicates: row/col compacted, values left in sorted order and cut to
    # the kept count (Q12), so (1, 2) carries 6.0 and (2, 3) 7.0 -- what the
    # reference loader leaves for dup_general.mtx too (ref_load.json)
    m = oracle.from_tuples(4, 4, np.array([0, 0, 1, 2]), np.array([1, 1, 2, 3]),
                           np.array([5.0, 6.0, 7.0, 8.0]), 4)
    assert m.row.tolist() == [0, 1, 2] and m.col.tolist() == [1, 2, 3]
    assert m.val.tolist() == [5.0, 6.0, 7.0]
    rec = [r for r in json.load(open(os.path.join(GOLD, "ref_load.json")))
           if r["file"] == "dup_general.mtx" and r["M"] == 4][0]
    assert (rec["row"], rec["col"], rec["val"]) == ([0, 1, 2, 3, 5], [1, 2, 0, 3, 5], [5.0, 6.0, 6.5, 7.0, 2.0])


def test_hybrid_fixtures_consistent():
    for name in ("GD01_c", "GD95_a", "lund_b", "bcsstm02", "small"):
        m = oracle.mtx_read(os.path.join(MTX, name + ".mtx"), 8)
        for V in (1, 8, 32):
            st, out = oracle.hybrid(m, 2, 8, V, 2, 0)
            if m.nnz == 0:
                continue
            assert out.nnz == m.nnz
            # final fields equal fresh evaluation of the final matrix
            cnt, _ = oracle.locate(out, 2, 8)
            assert st[4] == float((cnt != 0).sum()) and st[6] == float(cnt.sum())
            assert (st[9], st[11]) == oracle.venom_eval(out, 8, V)
            assert (st[8], st[10]) == oracle.venom_eval(m, 8, V)
            assert st[14] == st[12] + st[13]
            # output order = sorted at last reload (positions unique)
            key = out.row.astype(np.int64) * out.ncols + out.col
            assert len(np.unique(key)) == out.nnz


def test_retained_nnz_brute_force():
    import itertools
    rng = np.random.default_rng(11)
    for t in range(10):
        M = (4, 8)[t % 2]
        R = int(rng.integers(1, 6)) * M
        m = _rand(rng, R, rng.uniform(0.1, 0.8), M)
        for V in (1, 2, 4, 8):
            bm = L.dense_bitmaps(R, M, m.row.tolist(), m.col.tolist(), m.val.tolist())
            tot = 0
            for s in range(R // M):
                for b0 in range(0, R, V):
                    best = 0
                    for S in itertools.combinations(range(M), min(4, M)):
                        mask = sum(1 << (M - 1 - j) for j in S)
                        best = max(best, sum(min(2, bin(bm[i][s] & mask).count("1"))
                                             for i in range(b0, min(b0 + V, R))))
                    tot += best
            assert oracle.retained_nnz(m, V, 2, M) == tot


# ---------------------------------------------------------------- allpairs (M = 4)

def test_allpairs_tables_scored_vs_applied():
    # the double swap each score entry describes equals the applied pairtab
    # entry for codes 0..32 and differs for 33..35 (the reference mismatch)
    for i in range(36):
        lo, tk = L.exchange_of_pos(i)
        pt = L.AP_PAIRTAB[i]
        applied = (sorted([pt[0], pt[2]]), sorted([pt[1] - 4, pt[3] - 4]))
        assert (applied == (lo, tk)) == (i < 33), i


def _dense_m4(R, dens, seed):
    rng = np.random.default_rng(seed)
    A = rng.random((R, R)) < dens
    r, c = np.nonzero(A)
    return oracle.from_tuples(R, R, r.astype(np.int32), c.astype(np.int32),
                              rng.uniform(0.1, 1, r.size).astype(np.float32), 4), A


def test_allpairs_score_brute_force():
    # gains from relabelling the columns of a dense matrix and recounting
    N = 2
    for seed in range(4):
        m, A = _dense_m4(32, 0.35, seed)
        nseg = 8
        rng = np.random.default_rng(seed + 10)

        def cnt(D, sg):
            return int((D[:, sg * 4:sg * 4 + 4].sum(axis=1) > N).sum())
        for _ in range(6):
            s1, s2 = rng.choice(nseg, 2, replace=False)
            sc = oracle.swap_score_all(m, N, int(s1), int(s2))
            cols = list(range(s1 * 4, s1 * 4 + 4)) + list(range(s2 * 4, s2 * 4 + 4))
            for i in range(52):
                if i < 16:
                    b1, b2 = divmod(i, 4)
                    order = list(range(8))
                    order[b1], order[4 + b2] = order[4 + b2], order[b1]
                else:
                    order = L.AP_POS[i - 16]
                D = A.copy()
                D[:, cols] = A[:, [cols[k] for k in order]]
                assert sc[2 * i] == cnt(A, s1) - cnt(D, s1), (seed, s1, s2, i)
                assert sc[2 * i + 1] == cnt(A, s2) - cnt(D, s2), (seed, s1, s2, i)


@pytest.mark.parametrize("sched", [4, 5, 6, 7])
@pytest.mark.parametrize("clean", [False, True])
def test_allpairs_matches_literal_restatement(sched, clean):
    for seed in range(3):
        m, _ = _dense_m4(24 + 8 * seed, 0.3, 100 + seed)
        rows, cols, vals = list(m.row), list(m.col), list(m.val)
        want, over = L.allpairs_reorder(m.nrows, rows, cols, vals, 2, sched, clean=clean)
        got, ow = oracle.allpairs_reorder(m, 2, sched, clean=clean)
        assert list(got) == want and ow == over, (seed, list(got), want, ow, over)
        assert list(m.row) == rows and list(m.col) == cols
        if clean:
            assert ow == 0


def test_spmm_oracle_dense_and_threads():
    """The FP64 V:N:M product on a matrix the format keeps whole (each V x M
    block uses at most 4 columns, each row at most 2 of them) equals the dense
    product, and the OpenMP form is bit-identical to one thread."""
    rng = np.random.default_rng(11)
    R, M, V = 64, 8, 4
    A = np.zeros((R, R), np.float32)
    for blk in range(R // V):
        for seg in range(R // M):
            S = rng.choice(M, 4, replace=False)
            for r in range(blk * V, blk * V + V):
                for c in rng.choice(S, int(rng.integers(0, 3)), replace=False):
                    A[r, seg * M + c] = rng.uniform(0.5, 2.0) * rng.choice([-1.0, 1.0])
    r, c = np.nonzero(A)
    m = oracle.Matrix(R, R, r.astype(np.int32), c.astype(np.int32), A[r, c].astype(np.float32))
    B = rng.standard_normal((R, 48)).astype(np.float32)
    C1 = oracle.spmm_vnm_f64(m, V, M, B)
    np.testing.assert_allclose(C1, A.astype(np.float64) @ B.astype(np.float64), rtol=1e-12, atol=1e-12)
    C4 = oracle.spmm_vnm_f64(m, V, M, B, threads=4)
    assert np.array_equal(C1, C4)
    # a pruned case: the threaded form still matches one thread exactly
    big = (rng.random((256, 256)) < 0.2)
    r, c = np.nonzero(big)
    m2 = oracle.Matrix(256, 256, r.astype(np.int32), c.astype(np.int32),
                       rng.uniform(-1, 1, r.size).astype(np.float32))
    B2 = rng.standard_normal((256, 32)).astype(np.float32)
    assert np.array_equal(oracle.spmm_vnm_f64(m2, 16, 8, B2), oracle.spmm_vnm_f64(m2, 16, 8, B2, threads=4))


def _rect(rng, M):
    """A matrix with more padded rows than columns, its entries in the
    leading square part or anywhere."""
    C = int(rng.integers(2, 8)) * M
    R = C + int(rng.integers(1, 5)) * M
    lim = C if rng.random() < 0.6 else R
    mask = rng.random((lim, C)) < rng.uniform(0.05, 0.4)
    r, c = np.nonzero(mask)
    return oracle.from_tuples(R, C, r, c, rng.uniform(-1, 1, r.size), M)


def test_venom_rectangular_matches_literal():
    """new_id renames columns too, so with more padded rows than columns a
    column can land past the n // M segments, where generate_segmat raises
    IndexError (reordering.py:113-118).  The oracle refuses exactly where the
    literal restatement raises and agrees with it everywhere else."""
    rng = np.random.default_rng(12)
    raised = kept = 0
    for t in range(120):
        M = (4, 8)[t % 2]
        m = _rect(rng, M)
        V = int(rng.choice([1, 4, 16]))
        try:
            want = L.venom_reorder(m.nrows, m.ncols, m.row.tolist(), m.col.tolist(), M, V)
        except IndexError:
            want = None
        om = m.copy()
        try:
            got = oracle.venom_reorder(om, M, V)
        except RuntimeError:
            got = None
        assert (want is None) == (got is None), t
        if want is None:
            raised += 1
            continue
        kept += 1
        assert got.tolist() == want[0]
        assert om.row.tolist() == want[1] and om.col.tolist() == want[2]
    assert raised > 0 and kept > 0


def test_reload_refuses_rows_past_the_written_shape():
    """The reload writes coo_matrix(shape=(n, n)), n the padded column count
    (reordering.py:180); scipy refuses a row index >= n, and so does the
    oracle's reload and the hybrid that runs it."""
    from scipy.sparse import coo_matrix
    rng = np.random.default_rng(13)
    outcomes = set()
    for t in range(60):
        M = (4, 8)[t % 2]
        m = _rect(rng, M)
        n = (m.ncols + M - 1) // M * M
        try:
            coo_matrix((m.val, (m.row, m.col)), shape=(n, n))
            ok = True
        except ValueError:
            ok = False
        om = m.copy()
        try:
            oracle.reload(om, M)
            got = True
        except RuntimeError:
            got = False
        assert ok == got, t
        if ok:
            assert (om.nrows, om.ncols) == (n, n)
        outcomes.add(ok)
    assert outcomes == {True, False}
    # the hybrid refuses when a round's reload or VENOM step would, and on
    # success ends at the written shape
    hy = set()
    for t in range(80):
        M = (4, 8)[t % 2]
        m = _rect(rng, M)
        try:
            st, out = oracle.hybrid(m.copy(), 2, M, int(rng.choice([1, 4, 16])), 2, t % 4)
            n = (m.ncols + M - 1) // M * M
            assert (out.nrows, out.ncols) == (n, n)
            assert out.nnz == 0 or (out.row.max() < n and out.col.max() < n)
            hy.add(True)
        except RuntimeError:
            hy.add(False)
    assert hy == {True, False}


def test_venom_m16_matches_literal():
    """M = 16: the oracle's VENOM against the literal restatement of
    reordering.py (16-bit Gray-code hashes, the same literal 4 and 2)."""
    rng = np.random.default_rng(16)
    for t in range(8):
        R = int(rng.integers(2, 8)) * 16
        mask = rng.random((R, R)) < rng.uniform(0.02, 0.3)
        r, c = np.nonzero(mask)
        m = oracle.from_tuples(R, R, r, c, rng.uniform(-1, 1, r.size), 16)
        V = int(rng.choice([1, 4, 16]))
        want = L.venom_reorder(m.nrows, m.ncols, m.row.tolist(), m.col.tolist(), 16, V)
        om = m.copy()
        got = oracle.venom_reorder(om, 16, V)
        assert got.tolist() == want[0], t
        assert om.row.tolist() == want[1] and om.col.tolist() == want[2]


def test_config_goldens_pin_the_generators():
    """The oracle runs in tests/golden/configs.json were made on these seeded
    inputs: the generators still produce them bit for bit (the GPU tests
    compare against those runs)."""
    import vnm
    from golden.make_config_golden import CHUNGLU_SAMPLE, chunglu_sizes
    gold = json.load(open(os.path.join(GOLD, "configs.json")))

    def digest(m):
        h = hashlib.sha256()
        for a in (m.row, m.col, m.val):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()
    assert digest(vnm.gen_rmat(16, 524288, 2, M=8)) == gold["rmat65536_hybrid"]["input_sha256"]
    assert digest(vnm.gen_rmat(17, 1572864, 4, M=4)) == gold["rmat131072_sched0"]["input_sha256"]
    for i in CHUNGLU_SAMPLE:
        m = vnm.gen_chung_lu(chunglu_sizes()[i], 2.1, 12.0, 2000 + i, M=8)
        assert digest(m) == gold[f"chunglu{i}_sched0"]["input_sha256"]
    # a small sampled graph's oracle run is cheap: rerun it against the record
    m = vnm.gen_chung_lu(chunglu_sizes()[21], 2.1, 12.0, 2021, M=8)
    om = oracle.Matrix(m.nrows, m.ncols, m.row.copy(), m.col.copy(), m.val.copy())
    st, _ = oracle.nm_reorder(om, 2, 8, 0, 10)
    assert st.tolist() == gold["chunglu21_sched0"]["stats"]
    assert digest(om) == gold["chunglu21_sched0"]["sha256"]
