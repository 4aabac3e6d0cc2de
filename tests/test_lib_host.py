"""CPU tests of libvnm's host side: exports, loader, writer, generators.

The loader/writer are product host code (no GPU needed); they are checked
against the oracle, which is itself pinned to the reference loader.
"""
import os
import re

import numpy as np
import pytest

import oracle
import vnm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "vnm.h")


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()
    vnm.lib()


def test_exports_every_declared_symbol():
    names = set(re.findall(r"\b(vnm_[a-z0-9_]+)\s*\(", open(HDR).read()))
    L = vnm.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(vnm.EXPORTS) <= names


def test_no_device_fails_loudly():
    # this host has no GPU: context creation must report it, not fall back
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except ImportError:
        pass
    with pytest.raises(vnm.VnmError) as e:
        vnm.Context(0)
    assert e.value.code == vnm.E_NODEV


def _same(a, b):
    return (a.nrows, a.ncols, a.nnz) == (b.nrows, b.ncols, b.nnz) and \
        np.array_equal(a.row, b.row) and np.array_equal(a.col, b.col) and \
        np.array_equal(a.val, b.val)


@pytest.mark.parametrize("M", [4, 8])
def test_loader_matches_oracle_on_fixtures(M):
    for f in sorted(os.listdir(os.path.join(GOLD, "mtx"))):
        if f.endswith(".mtx"):
            p = os.path.join(GOLD, "mtx", f)
            assert _same(vnm.mtx_read(p, M), oracle.mtx_read(p, M)), f


def test_loader_matches_reference_loader_corners():
    # the library's loader (sequential and threaded) against the reference
    # loader's own output on the corner-case files (ref_load.json)
    import json
    for rec in json.load(open(os.path.join(GOLD, "ref_load.json"))):
        p = os.path.join(GOLD, rec["dir"], rec["file"])
        for threads in (1, 0):
            m = vnm.mtx_read(p, rec["M"], threads=threads)
            assert (m.nrows, m.ncols, m.nnz) == (rec["nrows"], rec["ncols"], rec["nnz"]), rec["file"]
            if "row" in rec:
                assert m.row.tolist() == rec["row"], rec["file"]
                assert m.col.tolist() == rec["col"], rec["file"]
                assert np.array_equal(m.val, np.array(rec["val"], np.float32)), rec["file"]


def test_tuples_with_duplicates_and_symmetry():
    rng = np.random.default_rng(3)
    for t in range(30):
        n = int(rng.integers(3, 40))
        k = int(rng.integers(0, 3 * n))
        r = rng.integers(0, n, k)
        c = rng.integers(0, n, k)
        v = rng.uniform(-1, 1, k).astype(np.float32)
        sym = bool(t % 2)
        for M in (4, 8):
            a = vnm.coo_from_tuples(n, n, r, c, v, M, symmetric=sym)
            b = oracle.from_tuples(n, n, r, c, v, M, symmetric=sym)
            assert _same(a, b)


def test_writer_matches_oracle_bytes(tmp_path):
    m = vnm.gen_er(100, 500, 9, -2.0, 2.0, 8)
    vnm.mtx_write(tmp_path / "a.mtx", m)
    oracle.mtx_write(tmp_path / "b.mtx", oracle.Matrix(m.nrows, m.ncols, m.row, m.col, m.val))
    assert (tmp_path / "a.mtx").read_bytes() == (tmp_path / "b.mtx").read_bytes()
    back = vnm.mtx_read(tmp_path / "a.mtx", 8)
    assert np.array_equal(back.row, m.row) and np.array_equal(back.col, m.col)


def test_loader_errors(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("not a banner\n1 1 1\n1 1 1\n")
    with pytest.raises(vnm.VnmError) as e:
        vnm.mtx_read(p, 8)
    assert e.value.code == vnm.E_FORMAT
    with pytest.raises(vnm.VnmError) as e:
        vnm.mtx_read(tmp_path / "missing.mtx", 8)
    assert e.value.code == vnm.E_IO
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n")
    with pytest.raises(vnm.VnmError) as e:
        vnm.mtx_read(p, 8)
    assert e.value.code == vnm.E_FORMAT
    p.write_text("%%MatrixMarket matrix coordinate complex general\n3 3 0\n")
    with pytest.raises(vnm.VnmError):
        vnm.mtx_read(p, 8)


def test_generators_deterministic_and_sized():
    a = vnm.gen_er(2048, round(0.01 * 2048 * 2048), 1)
    b = vnm.gen_er(2048, round(0.01 * 2048 * 2048), 1)
    assert a.nnz == 41943 and _same(a, b)
    assert a.val.min() >= -1 and a.val.max() < 1
    r = vnm.gen_rmat(10, 4000, 2)
    assert r.nnz == 8000 and r.nrows == 1024
    key = set(zip(r.row.tolist(), r.col.tolist()))
    assert all((c, rr) in key for rr, c in key)           # symmetric
    assert not any(rr == c for rr, c in key)              # no self loops
    cl = vnm.gen_chung_lu(1000, 2.1, 12, 2000)
    assert cl.nnz > 0 and cl.nrows == 1000
