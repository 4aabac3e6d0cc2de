"""GPU parity of the V:N:M packer and SpMM.

Retained nnz is integer work: bit-exact against the oracle.  The SpMM is
compared with the oracle's FP64 product of the same pruned matrix (FP32
values, FP32 B).  Tolerances (relative, normwise ||dC||_F / ||C||_F and
max|dC| / max|C|):
  A FP32, B FP32 : 1e-5   (FP32 accumulation only)
  A BF16, B BF16 : 1e-2   (bf16 rounding of both operands)
  A FP8,  B FP8  : 1e-1   (e4m3, per-block / per-128-row scales)
  A FP16, B FP32 : 2e-3   (fp16 rounding of the values: 2^-11 relative)
  A FP16, B BF16 : 1e-2   (bf16 rounding of B dominates)
The product is also bit-identical across runs and launch shapes: every
output value is one fixed sequence of FP32 operations (rows longer than a
work item are summed from per-chunk slots in item order; no atomics).
"""
import os

import numpy as np
import pytest

import oracle
import vnm

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    c = vnm.Context(0)
    yield c
    c.close()


def _om(m):
    return oracle.Matrix(m.nrows, m.ncols, m.row, m.col, m.val)


def _rand(rng, R, dens, M, zero_frac=0.05):
    mask = rng.random((R, R)) < dens
    r, c = np.nonzero(mask)
    v = rng.uniform(-1, 1, r.shape[0]).astype(np.float32)
    v[rng.random(r.shape[0]) < zero_frac] = 0.0
    return vnm.coo_from_tuples(R, R, r, c, v, M)


def test_retained_nnz_matches_oracle(ctx):
    rng = np.random.default_rng(41)
    cases = [(_rand(rng, int(rng.integers(1, 30)) * M, rng.uniform(0.02, 0.6), M), M)
             for M in (4, 8, 16) for _ in range(5)]
    cases.append((vnm.gen_er(2048, 41943, 1), 8))
    cases.append((vnm.gen_rmat(13, 60000, 3), 8))
    cases.append((vnm.gen_rmat(12, 40000, 4, M=4), 4))
    cases.append((vnm.mtx_read(os.path.join(GOLD, "mtx", "lund_b.mtx"), 8), 8))
    for m, M in cases:
        g = vnm.Graph(ctx, m, M)
        for V in (1, 8, 64, 128):
            assert g.retained_nnz(V) == oracle.retained_nnz(_om(m), V, 2, M), (M, V)


def _errs(C, ref):
    d = C.astype(np.float64) - ref
    fro = np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300)
    mx = np.abs(d).max() / max(np.abs(ref).max(), 1e-300)
    return fro, mx


def _e4m3_rows(B):
    """FP8 e4m3 B with one FP32 scale per 128 rows (absmax / 448)."""
    K = B.shape[0]
    nb = (K + 127) // 128
    scales = np.ones(nb, np.float32)
    Bs = np.empty_like(B)
    for i in range(nb):
        blk = B[i * 128:(i + 1) * 128]
        a = np.abs(blk).max()
        scales[i] = a / 448.0 if a > 0 else 1.0
        Bs[i * 128:(i + 1) * 128] = blk / scales[i]
    q = torch.from_numpy(Bs).to(torch.float8_e4m3fn)
    return q, scales


def _run(ctx, m, M, V, n, adt, bdt, seed=5, vec=None):
    g = vnm.Graph(ctx, m, M)
    p = g.pack(V, adt)
    if vec is not None:
        p.set_shape(bdt, n, vec)
    info = p.info()
    B = vnm.gen_normal(seed, (m.ncols, n))
    dev = torch.device("cuda:0")
    scales_d = None
    if bdt == vnm.DT_F32:
        Bd = torch.from_numpy(B).to(dev)
    elif bdt == vnm.DT_BF16:
        Bd = torch.from_numpy(B).to(torch.bfloat16).to(dev)
    else:
        q, sc = _e4m3_rows(B)
        Bd = q.to(dev)
        scales_d = torch.from_numpy(sc).to(dev)
    C = torch.full((m.nrows, n), float("nan"), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    p.spmm(Bd.data_ptr(), n, n, bdt, C.data_ptr(), n,
           scales_d.data_ptr() if scales_d is not None else None)
    torch.cuda.synchronize()
    ref = oracle.spmm_vnm_f64(_om(m), V, M, B)
    return C.cpu().numpy(), ref, info


@pytest.mark.parametrize("n", [64, 128, 256])
def test_spmm_fp32_exactish(ctx, n):
    rng = np.random.default_rng(n)
    for M in (4, 8, 16):
        m = _rand(rng, 40 * M, 0.05, M)
        C, ref, info = _run(ctx, m, M, 16, n, vnm.DT_F32, vnm.DT_F32)
        assert info.retained_nnz == oracle.retained_nnz(_om(m), 16, 2, M)
        fro, mx = _errs(C, ref)
        assert fro < 1e-5 and mx < 1e-5, (fro, mx)


def test_spmm_rmat_bf16_and_fp8(ctx):
    m = vnm.gen_rmat(13, 60000, 3)  # includes hub rows longer than one work item
    C, ref, info = _run(ctx, m, 8, 128, 256, vnm.DT_F32, vnm.DT_F32)
    assert max(_errs(C, ref)) < 1e-5
    C, ref, _ = _run(ctx, m, 8, 128, 256, vnm.DT_BF16, vnm.DT_BF16)
    fro, mx = _errs(C, ref)
    assert fro < 1e-2 and mx < 1e-2, (fro, mx)
    C, ref, _ = _run(ctx, m, 8, 128, 256, vnm.DT_FP8, vnm.DT_FP8)
    fro, mx = _errs(C, ref)
    assert fro < 1e-1 and mx < 1e-1, (fro, mx)


def test_spmm_fp16_values(ctx):
    # FP16 matrix values with FP32 and BF16 B, M = 8 and 4
    for mm, M, V in ((vnm.gen_rmat(13, 60000, 6), 8, 128), (vnm.gen_rmat(12, 40000, 7, M=4), 4, 64)):
        C, ref, _ = _run(ctx, mm, M, V, 128, vnm.DT_F16, vnm.DT_F32)
        fro, mx = _errs(C, ref)
        assert fro < 2e-3 and mx < 2e-3, (M, fro, mx)
        C, ref, _ = _run(ctx, mm, M, V, 256, vnm.DT_F16, vnm.DT_BF16)
        fro, mx = _errs(C, ref)
        assert fro < 1e-2 and mx < 1e-2, (M, fro, mx)


def test_spmm_m4_bf16(ctx):
    m = vnm.gen_rmat(12, 40000, 4, M=4)
    C, ref, _ = _run(ctx, m, 4, 64, 128, vnm.DT_BF16, vnm.DT_BF16)
    fro, mx = _errs(C, ref)
    assert fro < 1e-2 and mx < 1e-2, (fro, mx)


def test_spmm_empty_rows_written(ctx):
    # rows without kept entries must come out exactly 0 (C starts as NaN)
    m = vnm.coo_from_tuples(64, 64, np.array([0, 5]), np.array([3, 9]),
                            np.array([1.5, -2.0], np.float32), 8)
    C, ref, _ = _run(ctx, m, 8, 8, 64, vnm.DT_F32, vnm.DT_F32)
    assert np.isfinite(C).all()
    assert np.allclose(C, ref, rtol=1e-6, atol=1e-6)


TOL = {vnm.DT_F32: 1e-5, vnm.DT_BF16: 1e-2, vnm.DT_FP8: 1e-1}
SB = {vnm.DT_F32: 4, vnm.DT_BF16: 2, vnm.DT_FP8: 1}


def _vecs(dt, n):
    return [v for v in (1, 2, 4, 8, 16) if 2 <= v * SB[dt] <= 16 and n % (64 * v) == 0]


@pytest.mark.parametrize("dt,n", [(vnm.DT_F32, 256), (vnm.DT_F32, 128), (vnm.DT_BF16, 256),
                                  (vnm.DT_BF16, 128), (vnm.DT_FP8, 128), (vnm.DT_FP8, 1024)])
def test_spmm_every_shape(ctx, dt, n):
    # every launch shape the tuner may pick gives the oracle's product
    m = vnm.gen_rmat(12, 30000, 8)  # hub rows span several work items
    for v in _vecs(dt, n):
        C, ref, _ = _run(ctx, m, 8, 64, n, dt, dt, vec=v)
        fro, mx = _errs(C, ref)
        assert fro < TOL[dt] and mx < TOL[dt], (v, fro, mx)


def test_spmm_tuner_picks_a_usable_shape(ctx):
    m = vnm.gen_rmat(11, 10000, 9)
    g = vnm.Graph(ctx, m, 8)
    p = g.pack(32, vnm.DT_BF16)
    assert p.get_shape(vnm.DT_BF16, 128) == 0
    B = torch.zeros((m.ncols, 128), dtype=torch.bfloat16, device="cuda")
    C = torch.empty((m.nrows, 128), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    p.spmm(B.data_ptr(), 128, 128, vnm.DT_BF16, C.data_ptr(), 128)
    assert p.get_shape(vnm.DT_BF16, 128) in _vecs(vnm.DT_BF16, 128)
    with pytest.raises(vnm.VnmError):
        p.set_shape(vnm.DT_F32, 64, 2)  # 64 lanes x 2 columns do not tile 64 columns


def test_fp8_decode_every_code(ctx):
    # C = I * B with FP8 B holding every finite e4m3 code: the hardware
    # conversion must give the OCP e4m3 values torch decodes
    R, n = 512, 1024
    idx = np.arange(R, dtype=np.int32)
    m = vnm.coo_from_tuples(R, R, idx, idx, np.ones(R, np.float32), 8)
    g = vnm.Graph(ctx, m, 8)
    p = g.pack(1, vnm.DT_F32)  # V = 1: every row keeps its one entry
    assert p.info().retained_nnz == R
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F], np.uint8)
    Bq = np.resize(codes, R * n).reshape(R, n)
    Bd = torch.from_numpy(Bq).to("cuda").view(torch.float8_e4m3fn)
    sc = torch.ones((R + 127) // 128, dtype=torch.float32, device="cuda")
    C = torch.empty((R, n), dtype=torch.float32, device="cuda")
    ref = Bd.to(torch.float32)
    for v in _vecs(vnm.DT_FP8, n):
        p.set_shape(vnm.DT_FP8, n, v)
        torch.cuda.synchronize()  # inputs made on torch's stream are complete
        p.spmm(Bd.data_ptr(), n, n, vnm.DT_FP8, C.data_ptr(), n, sc.data_ptr())
        assert torch.equal(C, ref), v


def _bits(C):
    return C.view(torch.int32).cpu().numpy()


@pytest.mark.parametrize("adt,bdt,n", [(vnm.DT_F32, vnm.DT_F32, 256), (vnm.DT_BF16, vnm.DT_BF16, 256),
                                       (vnm.DT_FP8, vnm.DT_FP8, 128), (vnm.DT_F16, vnm.DT_F32, 128)])
def test_spmm_bit_identical_across_runs_and_shapes(ctx, adt, bdt, n):
    # hub rows of up to ~600 row entries: cut into many work items
    m = vnm.gen_rmat(14, 200000, 12)
    g = vnm.Graph(ctx, m, 8)
    p = g.pack(128, adt)
    B = vnm.gen_normal(13, (m.ncols, n))
    dev = torch.device("cuda:0")
    sc = None
    if bdt == vnm.DT_F32:
        Bd = torch.from_numpy(B).to(dev)
    elif bdt == vnm.DT_BF16:
        Bd = torch.from_numpy(B).to(torch.bfloat16).to(dev)
    else:
        q, s_ = _e4m3_rows(B)
        Bd, sc = q.to(dev), torch.from_numpy(s_).to(dev)
    first = None
    for v in _vecs(bdt, n):
        p.set_shape(bdt, n, v)
        for rep in range(2):
            C = torch.full((m.nrows, n), float("nan"), dtype=torch.float32, device=dev)
            torch.cuda.synchronize()  # inputs made on torch's stream are complete
            p.spmm(Bd.data_ptr(), n, n, bdt, C.data_ptr(), n, sc.data_ptr() if sc is not None else None)
            torch.cuda.synchronize()
            b = _bits(C)
            if first is None:
                first = b
                ref = oracle.spmm_vnm_f64(_om(m), 128, 8, B)
                fro, mx = _errs(C.cpu().numpy(), ref)
                tol = {vnm.DT_F32: 1e-5, vnm.DT_BF16: 1e-2, vnm.DT_FP8: 1e-1}[bdt]
                tol = 2e-3 if adt == vnm.DT_F16 else tol
                assert fro < tol and mx < tol, (fro, mx)
            else:
                assert np.array_equal(b, first), (v, rep)


def _bf16_run(p, m, n, B, vec=None, bias=None):
    if vec is not None:
        p.set_shape(vnm.DT_BF16, n, vec)
    Bd = torch.from_numpy(B).to(torch.bfloat16).cuda()
    C = torch.full((m.nrows, n), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    p.spmm(Bd.data_ptr(), n, n, vnm.DT_BF16, C.data_ptr(), n)
    torch.cuda.synchronize()
    return C.cpu().numpy()


@pytest.mark.parametrize("case", ["rmat_m8_v64", "rmat_m4_v64", "dense_m8_v16", "chunglu_v128"])
def test_spmm_sparse_mfma_part(ctx, case):
    """Blocks keeping >= dense_min values on v_smfmac_f32_16x16x64_bf16: the
    same pruned matrix (retained count unchanged), the FP64 oracle's product
    within the BF16 tolerance, bit-identical across runs and row-entry launch
    shapes, for thresholds that move none, some and all blocks."""
    rng = np.random.default_rng(7)
    if case == "rmat_m8_v64":
        m, M, V, n = vnm.gen_rmat(13, 60000, 31), 8, 64, 128
    elif case == "rmat_m4_v64":
        m, M, V, n = vnm.gen_rmat(12, 40000, 32, M=4), 4, 64, 256
    elif case == "dense_m8_v16":
        m, M, V, n = _rand(rng, 20 * 8, 0.3, 8), 8, 16, 64
    else:
        m, M, V, n = vnm.gen_chung_lu(4000, 2.1, 12.0, 33), 8, 128, 64
    g = vnm.Graph(ctx, m, M)
    B = vnm.gen_normal(34, (m.ncols, n))
    ref = oracle.spmm_vnm_f64(_om(m), V, M, B.astype(np.float32))
    retained = oracle.retained_nnz(_om(m), V, 2, M)
    # B rounded to BF16 as the kernels read it
    Bq = torch.from_numpy(B).to(torch.bfloat16).float().numpy()
    refq = oracle.spmm_vnm_f64(_om(m), V, M, Bq)
    seen_dense = False
    for dmin in (1, 4, 8, 1 << 20):
        p = g.pack(V, vnm.DT_BF16, dense_min=dmin)
        info = p.info()
        assert info.retained_nnz == retained
        if dmin == 1:
            assert info.dense_blocks == info.nblocks
        if dmin == 1 << 20:
            assert info.dense_blocks == 0
        seen_dense |= info.dense_blocks > 0
        first = None
        for v in _vecs(vnm.DT_BF16, n):
            for rep in range(2):
                C = _bf16_run(p, m, n, B, vec=v)
                if first is None:
                    first = C
                    fro, mx = _errs(C, ref)
                    assert fro < 1e-2 and mx < 1e-2, (dmin, fro, mx)
                    fro, mx = _errs(C, refq)  # only the matrix values' rounding left
                    assert fro < 5e-3 and mx < 5e-3, (dmin, fro, mx)
                else:
                    assert np.array_equal(C.view(np.int32), first.view(np.int32)), (dmin, v, rep)
        p.close()
    assert seen_dense


def test_spmm_sparse_mfma_gcn_bias_and_noalloc(ctx):
    """The GCN layer (bias in every written row, including rows only the 2:4
    part holds and rows cut into chunks) and the allocation-free entry on a
    matrix with a sparse-MFMA part."""
    m = vnm.gen_rmat(13, 80000, 35)  # hub rows cut into chunks
    g = vnm.Graph(ctx, m, 8)
    p = g.pack(64, vnm.DT_BF16, dense_min=4)
    info = p.info()
    assert info.dense_blocks > 0 and info.slots > 0
    K, Nout = 64, 128
    X = torch.from_numpy(vnm.gen_normal(36, (m.ncols, K))).to(torch.bfloat16).cuda()
    W = (torch.from_numpy(vnm.gen_normal(37, (K, Nout))) / 8.0).to(torch.bfloat16).cuda()
    b = torch.linspace(-0.5, 0.5, Nout, device="cuda")
    Y = torch.full((m.nrows, Nout), float("nan"), device="cuda")
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    p.gcn_layer(X.data_ptr(), K, K, W.data_ptr(), Nout, Y.data_ptr(), Nout, b.data_ptr())
    torch.cuda.synchronize()
    H = (X.double() @ W.double()).to(torch.bfloat16).float().cpu().numpy()
    ref = oracle.spmm_vnm_f64(_om(m), 64, 8, H) + b.double().cpu().numpy()
    fro, mx = _errs(Y.cpu().numpy(), ref)
    assert fro < 1e-2 and mx < 1e-2, (fro, mx)
    n = 128
    B = vnm.gen_normal(38, (m.ncols, n))
    C1 = _bf16_run(p, m, n, B)
    Bd = torch.from_numpy(B).to(torch.bfloat16).cuda()
    C2 = torch.full((m.nrows, n), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    p.spmm_noalloc(Bd.data_ptr(), n, n, vnm.DT_BF16, C2.data_ptr(), n)
    torch.cuda.synchronize()
    assert np.array_equal(C1.view(np.int32), C2.cpu().numpy().view(np.int32))


def test_spmm_sparse_mfma_rejects(ctx):
    m = vnm.gen_rmat(10, 5000, 39)
    g = vnm.Graph(ctx, m, 8)
    with pytest.raises(vnm.VnmError):
        g.pack(64, vnm.DT_F32, dense_min=4)      # BF16 only
    with pytest.raises(vnm.VnmError):
        g.pack(24, vnm.DT_BF16, dense_min=4)     # V a multiple of 16
    p = g.pack(64, vnm.DT_BF16, dense_min=2)
    assert p.info().dense_blocks > 0
    B = torch.zeros((m.ncols, 64), dtype=torch.float32, device="cuda")
    C = torch.zeros((m.nrows, 64), dtype=torch.float32, device="cuda")
    with pytest.raises(vnm.VnmError):
        p.spmm(B.data_ptr(), 64, 64, vnm.DT_F32, C.data_ptr(), 64)  # BF16 B only
    Bb = torch.zeros((m.ncols, 64), dtype=torch.bfloat16, device="cuda")
    with pytest.raises(vnm.VnmError):
        vnm.Batch(ctx, [p], [Bb.data_ptr()], [64], 64, vnm.DT_BF16, [C.data_ptr()], [64])
