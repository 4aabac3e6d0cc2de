"""GPU parity of the batched SpMM (HIP graph replays) and the GCN layer.

The batch must give, on every replay of its graph, bit for bit the products
of one vnm_spmm per matrix, and the FP64 oracle's within the dtype tolerance (FP32 1e-5,
BF16 1e-2, normwise and max-relative); replays are separated by other eager
work on the same stream and host synchronisations, and the outputs are
poisoned with NaN before each replay so every row must be rewritten.

The GCN layer Y = A_vnm (X W) + b is checked against the FP64 oracle's
product of the same pruned matrix with H = X W rounded to BF16 the way the
layer stores it (tolerance 1e-2: BF16 rounding of X, W and H).
"""
import numpy as np
import pytest

import oracle
import vnm

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
N, M = 2, 8


@pytest.fixture(scope="module")
def ctx():
    c = vnm.Context(0)
    yield c
    c.close()


def _om(m):
    return oracle.Matrix(m.nrows, m.ncols, m.row.copy(), m.col.copy(), m.val.copy())


def _errs(C, ref):
    d = C.astype(np.float64) - ref
    return (np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300),
            np.abs(d).max() / max(np.abs(ref).max(), 1e-300))


def _graphs(seed0, count):
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(500), np.log(6000))))
        out.append(vnm.gen_chung_lu(n, 2.1, 12.0, seed0 + i))
    return out


@pytest.mark.parametrize("bdt,n", [(vnm.DT_BF16, 256), (vnm.DT_F32, 128)])
def test_batch_replays_match_single_and_oracle(ctx, bdt, n):
    tol = 1e-2 if bdt == vnm.DT_BF16 else 1e-5
    coos = _graphs(300 + n, 6)
    graphs = [vnm.Graph(ctx, m, M) for m in coos]
    packed = [g.pack(64, bdt) for g in graphs]
    tdt = torch.bfloat16 if bdt == vnm.DT_BF16 else torch.float32
    Bs = [torch.from_numpy(vnm.gen_normal(7 + i, (m.ncols, n))).to(tdt).cuda() for i, m in enumerate(coos)]
    Cs = [torch.empty((m.nrows, n), dtype=torch.float32, device="cuda") for m in coos]
    bias = [torch.linspace(-1, 1, n, device="cuda") if i % 2 else None for i in range(len(coos))]
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    batch = vnm.Batch(ctx, packed, [b.data_ptr() for b in Bs], [n] * len(coos), n, bdt,
                      [c.data_ptr() for c in Cs], [n] * len(coos),
                      bias=[b.data_ptr() if b is not None else 0 for b in bias])
    info = batch.info()
    assert info["cols_per_lane"] in (1, 2, 4, 8) and info["units"] > 0
    # eager references: one vnm_spmm per matrix (+ bias on the host side)
    single = []
    for p, B, bi in zip(packed, Bs, bias):
        C1 = torch.empty((p.info().nrows, n), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()  # inputs made on torch's stream are complete
        p.spmm(B.data_ptr(), n, n, bdt, C1.data_ptr(), n)
        torch.cuda.synchronize()
        single.append(C1.cpu().numpy() + (bi.cpu().numpy() if bi is not None else 0))
    refs = [oracle.spmm_vnm_f64(_om(m), 64, M, B.float().cpu().numpy().astype(np.float32))
            + (bi.double().cpu().numpy() if bi is not None else 0) for m, B, bi in zip(coos, Bs, bias)]
    scratch = torch.zeros(1 << 20, device="cuda")
    for rep in range(4):
        for C in Cs:
            C.fill_(float("nan"))
        scratch.add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
        batch.run()
        torch.cuda.synchronize()
        for i, C in enumerate(Cs):
            got = C.cpu().numpy()
            assert np.isfinite(got).all(), (rep, i)
            fro, mx = _errs(got, refs[i])
            assert fro < tol and mx < tol, (rep, i, fro, mx)
            # the batch runs the single kernel's per-item code: same bits
            assert np.array_equal(got.view(np.int32), single[i].view(np.int32)), (rep, i)
    batch.close()


# K 48 and 272 end on a half step of the 32-deep MFMA (272: in the third
# 128-deep LDS stage); ldx > K reads X through a row pitch
@pytest.mark.parametrize("K,ldx", [(96, 96), (48, 52), (272, 272), (256, 260)])
def test_gcn_layer_matches_oracle(ctx, K, ldx):
    m = vnm.gen_chung_lu(3000, 2.1, 12.0, 77)
    g = vnm.Graph(ctx, m, M)
    p = g.pack(64, vnm.DT_BF16)
    Nout = 128
    Xfull = torch.from_numpy(vnm.gen_normal(1, (m.ncols, ldx))).to(torch.bfloat16).cuda()
    X = Xfull[:, :K]
    W = (torch.from_numpy(vnm.gen_normal(2, (K, Nout))) / np.sqrt(K)).to(torch.bfloat16).cuda()
    b = torch.linspace(-0.5, 0.5, Nout, device="cuda")
    Y = torch.full((m.nrows, Nout), float("nan"), device="cuda")
    torch.cuda.synchronize()  # inputs made on torch's stream are complete
    p.gcn_layer(Xfull.data_ptr(), ldx, K, W.data_ptr(), Nout, Y.data_ptr(), Nout, b.data_ptr())
    torch.cuda.synchronize()
    H = (X.double() @ W.double()).to(torch.bfloat16).float().cpu().numpy()
    ref = oracle.spmm_vnm_f64(_om(m), 64, M, H) + b.double().cpu().numpy()
    fro, mx = _errs(Y.cpu().numpy(), ref)
    assert fro < 1e-2 and mx < 1e-2, (fro, mx)


def test_gcn_layer_rejects_bad_shapes(ctx):
    m = vnm.gen_chung_lu(600, 2.1, 12.0, 5)
    p = vnm.Graph(ctx, m, M).pack(64, vnm.DT_BF16)
    X = torch.zeros((m.ncols, 40), dtype=torch.bfloat16, device="cuda")
    W = torch.zeros((40, 64), dtype=torch.bfloat16, device="cuda")
    Y = torch.zeros((m.nrows, 64), device="cuda")
    with pytest.raises(vnm.VnmError):
        p.gcn_layer(X.data_ptr(), 40, 40, W.data_ptr(), 64, Y.data_ptr(), 64)  # K % 16 != 0
