"""GPU parity of VENOM, the hybrid driver, find_best_2toX and the CLI.

The oracle's VENOM is pinned to reordering.py itself (tests/golden/
venom_ref.json); here the device path is compared with the oracle
bit-for-bit and with the script's own recorded outputs.
"""
import json
import os
import subprocess

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


def _om(m):
    return oracle.Matrix(m.nrows, m.ncols, m.row, m.col, m.val)


def _rand(rng, R, dens, M, sym=False):
    mask = rng.random((R, R)) < dens
    if sym:
        mask |= mask.T
    r, c = np.nonzero(mask)
    v = rng.uniform(-1, 1, r.shape[0]).astype(np.float32)
    return vnm.coo_from_tuples(R, R, r, c, v, M)


def test_venom_matches_reference_script(ctx):
    for c in json.load(open(os.path.join(GOLD, "venom_ref.json"))):
        m = vnm.HostCoo(c["nrows"], c["ncols"], c["row"], c["col"], c["val"])
        g = vnm.Graph(ctx, m, c["M"])
        assert list(g.venom_eval(c["V"])) == c["evalonly"], c["name"]
        assert g.venom_reorder(c["V"], 10) == c["output"], (c["name"], c["V"])
        out = g.download()
        pos = sorted(set(zip(out.row.tolist(), out.col.tolist())))
        assert [list(p) for p in pos] == c["out_pos"]


def test_venom_matches_oracle_random(ctx):
    rng = np.random.default_rng(21)
    for t in range(16):
        M = (4, 8)[t % 2]
        R = int(rng.integers(2, 40)) * M
        m = _rand(rng, R, rng.uniform(0.01, 0.4), M, sym=t % 3 == 0)
        for V in (1, 4, 16, 64):
            g = vnm.Graph(ctx, m, M)
            st = g.venom_reorder(V, 10)
            out = g.download()
            om = _om(m).copy()
            ost = oracle.venom_reorder(om, M, V, 10)
            oracle.reload(om, M)
            assert st == ost.tolist()
            assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col)
            assert np.array_equal(out.val, om.val)


def test_venom_rmat_medium(ctx):
    m = vnm.gen_rmat(13, 60000, 3)
    for V in (32, 128):
        g = vnm.Graph(ctx, m, 8)
        assert g.venom_eval(V) == oracle.venom_eval(_om(m), 8, V)
        st = g.venom_reorder(V, 10)
        om = _om(m).copy()
        ost = oracle.venom_reorder(om, 8, V, 10)
        oracle.reload(om, 8)
        out = g.download()
        assert st == ost.tolist()
        assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col)


def _perm_maps_positions(g, m, out):
    """Every relabel renames rows and columns alike and the reload only
    re-sorts, so the final positions are the input's mapped through the
    composed relabel."""
    perm = g.perm()
    assert perm.shape == (m.nrows,)
    want = np.unique(perm[m.row].astype(np.int64) << 32 | perm[m.col].astype(np.int64))
    got = np.unique(out.row.astype(np.int64) << 32 | out.col.astype(np.int64))
    assert np.array_equal(want, got)


def _hyb(ctx, m, M, V, rounds, sched):
    g = vnm.Graph(ctx, m, M)
    st = g.hybrid(2, V, rounds, sched).as_tuple()
    out = g.download()
    ost, oout = oracle.hybrid(_om(m), 2, M, V, rounds, sched)
    assert st[:15] == ost, (st, ost)
    assert np.array_equal(out.row, oout.row) and np.array_equal(out.col, oout.col)
    assert np.array_equal(out.val, oout.val)
    if m.nnz:
        _perm_maps_positions(g, m, out)
    return st


def test_hybrid_er_2048_config(ctx):
    # 2,048 x 2,048 ER at 1% (41,943 entries, U(-1,1), seed 1), V:N:M = 64:2:8,
    # 10 rounds, schedule 0
    st = _hyb(ctx, vnm.gen_er(2048, 41943, 1), 8, 64, 10, 0)
    assert st[3] == 22.0


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_hybrid_fixtures(ctx, sched):
    for f in ("GD01_c", "GD95_a", "lund_b", "bcsstm02"):
        m = vnm.mtx_read(os.path.join(GOLD, "mtx", f + ".mtx"), 8)
        for V in (1, 8, 32):
            _hyb(ctx, m, 8, V, 3, sched)


def test_hybrid_random_and_m4(ctx):
    rng = np.random.default_rng(33)
    for t in range(6):
        M = (4, 8)[t % 2]
        R = int(rng.integers(4, 40)) * M
        m = _rand(rng, R, rng.uniform(0.02, 0.3), M, sym=t % 2 == 0)
        _hyb(ctx, m, M, int(rng.choice([1, 4, 16])), 2, t % 4)


def test_find_best_sched(ctx):
    for m in (vnm.gen_er(2048, 41943, 1), vnm.gen_rmat(12, 32768, 11),
              vnm.mtx_read(os.path.join(GOLD, "mtx", "lund_b.mtx"), 8)):
        g = vnm.Graph(ctx, m, 8)
        assert g.find_best_sched(2, 4) == oracle.find_best_sched(_om(m), 2, 8, 4)


def test_cli_matches_oracle_cli(tmp_path):
    for f in ("lund_b", "GD01_c"):
        inp = os.path.join(GOLD, "mtx", f + ".mtx")
        a = subprocess.run([vnm.CLI_PATH, "--mtxfile", inp, "--outmtxfile", str(tmp_path / "a.mtx"),
                            "--maxiter", "3", "--v", "8", "--n", "16", "--sched", "0"],
                           capture_output=True, text=True, check=True).stdout.strip().split(", ")
        b = subprocess.run([oracle.CLI_PATH, "--mtxfile", inp, "--outmtxfile", str(tmp_path / "b.mtx"),
                            "--maxiter", "3", "--v", "8", "--n", "16", "--sched", "0"],
                           capture_output=True, text=True, check=True).stdout.strip().split(", ")
        assert a[:15] == b[:15]
        assert (tmp_path / "a.mtx").read_bytes() == (tmp_path / "b.mtx").read_bytes()


def test_kat_temp_mtx_venom(ctx):
    m = vnm.mtx_read(os.path.join(GOLD, "mtx", "temp.mtx"), 8)
    g = vnm.Graph(ctx, m, 8)
    for V in (1, 8, 16, 32):
        assert g.venom_eval(V) == (0, 0)


def _rect(rng, M):
    """More padded rows than columns, entries in the leading square part or
    anywhere."""
    C = int(rng.integers(2, 8)) * M
    R = C + int(rng.integers(1, 5)) * M
    lim = C if rng.random() < 0.6 else R
    mask = rng.random((lim, C)) < rng.uniform(0.05, 0.4)
    r, c = np.nonzero(mask)
    return vnm.coo_from_tuples(R, C, r, c, rng.uniform(-1, 1, r.size).astype(np.float32), M)


def test_rectangular_hybrid_venom_and_pack(ctx):
    """The row permutation renames columns too, so on a matrix with more
    padded rows than columns VENOM, the reload to the (n, n) shape and the
    packer can meet entries past the padded column count, where the
    reference's Python raises.  The device refuses exactly where the oracle
    does; elsewhere its results, the shape the reload leaves and the greedy
    rounds captured before the rows shrank are the oracle's."""
    rng = np.random.default_rng(44)
    hyb, packs = set(), set()
    for t in range(48):
        M = (4, 8)[t % 2]
        m = _rect(rng, M)
        V = int(rng.choice([1, 4, 16]))
        sched = t % 4
        print("case", t, M, m.nrows, m.ncols, V, sched, flush=True)
        g = vnm.Graph(ctx, m, M)
        try:
            st, err = g.hybrid(2, V, 2, sched).as_tuple(), None
        except vnm.VnmError as e:
            st, err = None, e
        try:
            (ost, oout), oerr = oracle.hybrid(_om(m), 2, M, V, 2, sched), None
        except RuntimeError as e:
            oerr = e
        assert (err is None) == (oerr is None), (t, err, oerr)
        if err is None:
            out = g.download()
            assert st[:15] == ost, (t, st, ost)
            assert (g.nrows, g.ncols) == (oout.nrows, oout.ncols)
            assert np.array_equal(out.row, oout.row) and np.array_equal(out.col, oout.col)
            assert np.array_equal(out.val, oout.val)
            _perm_maps_positions(g, m, out)
            # a clone keeps the composed relabel over the ids at creation
            c = g.clone()
            assert np.array_equal(c.perm(), g.perm()) and (c.nrows, c.ncols) == (g.nrows, g.ncols)
            c.close()
        hyb.add(err is None)
        g.close()
        # one N:M run, then the packer
        g2 = vnm.Graph(ctx, m, M)
        g2.nm_reorder(2, sched)
        om = _om(m).copy()
        oracle.nm_reorder(om, 2, M, sched)
        out2 = g2.download()
        assert np.array_equal(out2.row, om.row) and np.array_equal(out2.col, om.col)
        over = om.nnz > 0 and int(om.col.max()) >= (m.ncols + M - 1) // M * M
        try:
            got = g2.retained_nnz(V)
        except vnm.VnmError:
            got = None
        assert (got is None) == over, t
        if got is not None:
            assert got == oracle.retained_nnz(om, V, 2, M)
        packs.add(over)
        g2.close()
    assert hyb == {True, False} and packs == {True, False}


def test_venom_m16_matches_oracle(ctx):
    """M = 16: 16-bit run bitmaps, the strict flag above them and row tokens
    scaled by 2^17; counts, stats and positions are the oracle's."""
    rng = np.random.default_rng(61)
    for t in range(8):
        R = int(rng.integers(2, 12)) * 16
        m = _rand(rng, R, rng.uniform(0.02, 0.3), 16, sym=t % 2 == 0)
        for V in (1, 4, 16, 64):
            g = vnm.Graph(ctx, m, 16)
            assert g.venom_eval(V) == oracle.venom_eval(_om(m), 16, V), (t, V)
            st = g.venom_reorder(V, 10)
            out = g.download()
            om = _om(m).copy()
            ost = oracle.venom_reorder(om, 16, V, 10)
            oracle.reload(om, 16)
            assert st == ost.tolist(), (t, V)
            assert np.array_equal(out.row, om.row) and np.array_equal(out.col, om.col)
            assert np.array_equal(out.val, om.val)
            g.close()
    # 32-bit row tokens bound the segment count at 16,382
    n = 16 * 16383
    g = vnm.Graph(ctx, vnm.coo_from_tuples(n, n, [0], [1], np.ones(1, np.float32), 16), 16)
    with pytest.raises(vnm.VnmError):
        g.venom_reorder(4, 10)
    g.close()
