"""The threaded MatrixMarket reader and writer against the sequential ones.

threads=1 runs the reference's loops (fscanf-equivalent strtol/strtof scan
and std::sort, mtx/readMtx.hpp:84-147,56-82; one fprintf per entry,
mtx/Mtx.h:412-435), which the loader tests pin to the reference loader.
Other thread counts parse whitespace-aligned parts of the text with exact
fast paths for plain decimal tokens, sort in parts and merge, and format
with an exact integer %.2f; every result and every error message must be
the sequential one.  Inputs are over 1 MiB so the threaded scan is taken.
"""
import os

import numpy as np
import pytest

import vnm

HEADER = "%%MatrixMarket matrix coordinate {field} {sym}\n"


def _read(path, M, threads):
    try:
        return vnm.mtx_read(path, M, threads), None
    except vnm.VnmError as e:
        return None, str(e)


def _same(path, M=8):
    a, ea = _read(path, M, 1)
    for th in (3, 8):
        b, eb = _read(path, M, th)
        assert ea == eb, (th, ea, eb)
        if a is None:
            continue
        assert (a.nrows, a.ncols) == (b.nrows, b.ncols)
        assert np.array_equal(a.row, b.row) and np.array_equal(a.col, b.col)
        assert np.array_equal(a.val.view(np.uint32), b.val.view(np.uint32))
    return a, ea


def _value_tokens(rng, n):
    """Value texts of every shape the scan meets: plain decimals (fast path),
    long mantissas and exponents (strtof), signs, bare points, specials."""
    fixed = ["0.41", "-0.0", "+3.25", "1.", ".5", "007", "0.1234567", "16777216", "16777217",
             "12345678.9", "0.00000000001", "1e-40", "3.4e39", "-3.4e39", "inf", "-inf", "nan",
             "0x1.8p1", "1E5", "-2.5e-3", "0.125", "99999999999999999999", "4.9406564584124654e-324"]
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 9, n)
    out = []
    for i in range(n):
        k = i % 7
        if k == 0:
            out.append("%.2f" % v[i])
        elif k == 1:
            out.append("%.9g" % v[i])
        elif k == 2:
            out.append("%.3e" % v[i])
        elif k == 3:
            out.append(fixed[(i // 7) % len(fixed)])
        elif k == 4:
            out.append(str(int(rng.integers(-1000, 1000))))
        elif k == 5:
            out.append("%.7f" % (v[i] % 1.0))
        else:
            out.append("%.1f" % v[i])
    return out


def _write_text(path, field, sym, nr, nc, entries, nv=None, seps=None):
    with open(path, "w", newline="") as f:
        f.write(HEADER.format(field=field, sym=sym))
        f.write("% a comment line\n")
        f.write("%d %d %d\n" % (nr, nc, len(entries) if nv is None else nv))
        for i, e in enumerate(entries):
            sep = seps[i % len(seps)] if seps else (" ", "\n")
            f.write(sep[0].join(e) + sep[1])


def _entries(rng, n, nr, nc, field, unique=True):
    if unique:
        flat = rng.choice(nr * nc, size=n, replace=False)
        r, c = flat // nc + 1, flat % nc + 1
    else:
        r, c = rng.integers(1, nr + 1, n), rng.integers(1, nc + 1, n)
    if field == "real":
        vals = _value_tokens(rng, n)
    elif field == "integer":
        vals = [str(int(x)) for x in rng.integers(-10 ** 6, 10 ** 6, n)]
    else:
        vals = None
    return [(str(a), str(b)) + ((vals[i],) if vals else ()) for i, (a, b) in enumerate(zip(r, c))]


@pytest.mark.parametrize("field,sym,unique", [("real", "general", True), ("real", "symmetric", True),
                                              ("integer", "general", True), ("pattern", "symmetric", True),
                                              ("real", "general", False)])
def test_reader_threads_match_sequential(tmp_path, field, sym, unique):
    rng = np.random.default_rng(hash((field, sym, unique)) % 2 ** 32)
    nr, nc = 5003, 4999
    ent = _entries(rng, 140000 if field == "pattern" else 90000, nr, nc, field, unique)
    # separators: spaces, tabs, CRLF, entries split over lines, blank lines
    seps = [(" ", "\n"), ("\t", "\n"), ("  ", "\r\n"), ("\n", "\n"), (" ", "\n\n")]
    path = str(tmp_path / "a.mtx")
    _write_text(path, field, sym, nr, nc, ent, seps=seps)
    assert os.path.getsize(path) > (1 << 20)
    a, err = _same(path)
    assert err is None and a.nnz > 0
    if not unique:
        assert a.nnz < len(ent)  # repeated positions were removed


def _big_entries(rng, n=80000, nr=3000):
    return _entries(rng, n, nr, nr, "real")


@pytest.mark.parametrize("case", ["bad_row", "range_then_bad", "bad_then_range", "range_and_bad_value",
                                  "truncated_value", "truncated_col", "unclean", "long_token",
                                  "short", "extra_garbage", "col_out_of_range"])
def test_reader_threads_same_errors(tmp_path, case):
    rng = np.random.default_rng(7)
    nr = 3000
    ent = [list(e) for e in _big_entries(rng, nr=nr)]
    n = len(ent)
    nv = None
    tail = ""
    if case == "bad_row":
        ent[50000][0] = "abc"
    elif case == "range_then_bad":
        ent[30000][0] = str(nr + 1)
        ent[60000][1] = "x1"
    elif case == "bad_then_range":
        ent[30000][2] = "zz"
        ent[60000][0] = "0"
    elif case == "range_and_bad_value":
        ent[40000][0] = "-5"
        ent[40000][2] = "q"
    elif case == "truncated_value":
        ent[-1] = ent[-1][:2]
    elif case == "truncated_col":
        ent[-1] = ent[-1][:1]
    elif case == "unclean":  # the sequential scan splits "12abc" into 12 and "abc"
        ent[20000][1] = "12abc"
    elif case == "long_token":
        ent[20000][2] = "0." + "1" * 80
    elif case == "short":
        nv = n + 10
    elif case == "extra_garbage":
        nv = n - 5
        tail = "garbage that is never read\n"
    elif case == "col_out_of_range":
        ent[45000][1] = str(10 ** 12)
    path = str(tmp_path / "b.mtx")
    _write_text(path, "real", "general", nr, nr, [tuple(e) for e in ent], nv=nv)
    with open(path, "a") as f:
        f.write(tail)
    a, err = _same(path)
    expect_ok = case in ("short", "extra_garbage", "long_token")
    assert (err is None) == expect_ok, err
    if case == "short":
        assert a.nnz == n
    if case == "extra_garbage":
        assert a.nnz == n - 5


def test_writer_threads_match_fprintf(tmp_path):
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2 ** 32, 600000, dtype=np.uint64).astype(np.uint32)
    special = np.array([0.0, -0.0, 0.125, 0.375, 0.625, 0.875, 1.125, 2.5, -0.125, 0.005, 0.015, 0.025,
                        1e-45, -1e-45, 1.17549435e-38, 2147483520.0, 2147483648.0, -2147483648.0,
                        3.4028235e38, np.inf, -np.inf, np.nan, 0.994999, 0.995, 0.9950001, 99.995,
                        -99.995, 1e-3, 4.999999e-3, 5.000001e-3], np.float32)
    ties = (np.arange(-4000, 4000) / 8.0).astype(np.float32)  # exact halves of a cent and more
    cents = (rng.integers(-10 ** 6, 10 ** 6, 200000) / 200.0).astype(np.float32)
    logs = (rng.standard_normal(200000) * 10.0 ** rng.integers(-12, 12, 200000)).astype(np.float32)
    vals = np.concatenate([bits.view(np.float32), special, ties, cents, logs])
    n = vals.size
    R = 1 << 20
    row = rng.integers(0, R, n).astype(np.int32)
    col = rng.integers(0, R, n).astype(np.int32)
    coo = vnm.coo_from_tuples(R, R, row, col, vals, 8)
    # coo_from_tuples sorts and deduplicates; both writers get the same COO
    outs = []
    for th in (1, 2, 8):
        p = tmp_path / ("w%d.mtx" % th)
        vnm.mtx_write(str(p), coo, th)
        outs.append(p.read_bytes())
    assert outs[0] == outs[1] == outs[2]
    small = vnm.coo_from_tuples(16, 16, [0, 3], [1, 2], [0.125, -0.0], 8)
    a, b = tmp_path / "s1.mtx", tmp_path / "s8.mtx"
    vnm.mtx_write(str(a), small, 1)
    vnm.mtx_write(str(b), small, 8)
    assert a.read_bytes() == b.read_bytes()


def test_read_write_round_trip_threads(tmp_path):
    m = vnm.gen_rmat(14, 120000, 9)
    p1, p8 = tmp_path / "r1.mtx", tmp_path / "r8.mtx"
    vnm.mtx_write(str(p1), m, 1)
    vnm.mtx_write(str(p8), m, 8)
    assert p1.read_bytes() == p8.read_bytes()
    a, err = _same(str(p8))
    assert err is None and a.nnz == m.nnz
    assert np.array_equal(a.row, m.row) and np.array_equal(a.col, m.col)


def test_binary_cache(tmp_path):
    """The cache file is named as the reference names it, serves the same COO
    for any M, and is ignored once the .mtx changes or the cache is damaged."""
    import shutil
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mtx", "lund_b.mtx")
    path = str(tmp_path / "lund_b.mtx")
    shutil.copy(src, path)
    sym = "symmetric" in open(path).readline().lower()
    cache = tmp_path / (".lund_b.mtx.%s.nosl.bin" % ("ud" if sym else "d"))
    for M, want_cache in ((8, False), (8, True), (4, True)):
        got, used = vnm.mtx_read_cached(path, M)
        ref = vnm.mtx_read(path, M)
        assert used == want_cache and cache.exists()
        assert (got.nrows, got.ncols) == (ref.nrows, ref.ncols)
        assert np.array_equal(got.row, ref.row) and np.array_equal(got.col, ref.col)
        assert np.array_equal(got.val.view(np.uint32), ref.val.view(np.uint32))
    # the source changes: the cache is rebuilt from the new text
    with open(path, "a") as f:
        f.write("\n")
    os.utime(path, ns=(0, 1_000_000_000))
    got, used = vnm.mtx_read_cached(path, 8)
    assert not used
    got, used = vnm.mtx_read_cached(path, 8)
    assert used
    # a damaged cache is ignored and replaced: cut short, or of the right
    # size with two entries out of order
    data = bytearray(cache.read_bytes())
    cache.write_bytes(bytes(data[: len(data) // 2]))
    got, used = vnm.mtx_read_cached(path, 8)
    ref = vnm.mtx_read(path, 8)
    assert not used and np.array_equal(got.row, ref.row) and np.array_equal(got.col, ref.col)
    got, used = vnm.mtx_read_cached(path, 8)
    assert used
    data = bytearray(cache.read_bytes())
    nnz = (len(data) - 64) // 12
    rows = np.frombuffer(bytes(data[64:64 + 4 * nnz]), np.int32).copy()
    k = int(np.nonzero(rows[1:] > rows[:-1])[0][0])  # rows[k] < rows[k + 1]
    rows[k], rows[k + 1] = rows[k + 1], rows[k]
    data[64:64 + 4 * nnz] = rows.tobytes()
    cache.write_bytes(bytes(data))
    got, used = vnm.mtx_read_cached(path, 8)
    assert not used and np.array_equal(got.row, ref.row) and np.array_equal(got.col, ref.col)
