"""The allocation-free SpMM entry point (vnm_spmm_noalloc) against vnm_spmm.

Two calls in a row on one context, the second with a smaller matrix, fewer
columns and another dtype: each must give vnm_spmm's product bit for bit
(and the FP64 oracle's within the dtype tolerance: FP32 1e-5, BF16 1e-2),
must leave the device's free memory unchanged, and a product needing more
partial-sum slots than the static array holds is refused with VNM_E_ARG.
"""
import numpy as np
import pytest

import oracle
import vnm

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _om(m):
    return oracle.Matrix(m.nrows, m.ncols, m.row.copy(), m.col.copy(), m.val.copy())


def _errs(C, ref):
    d = C.astype(np.float64) - ref
    return (np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300),
            np.abs(d).max() / max(np.abs(ref).max(), 1e-300))


def test_noalloc_matches_spmm_two_calls():
    ctx = vnm.Context(0)
    cases = [  # (matrix, M, V, n, dtype, tolerance); the second is smaller
        (vnm.gen_rmat(15, 400000, 21), 8, 128, 256, vnm.DT_F32, 1e-5),
        (vnm.gen_chung_lu(3000, 2.1, 12.0, 22), 8, 64, 64, vnm.DT_BF16, 1e-2),
    ]
    prepared = []
    for m, M, V, n, dt, tol in cases:
        g = vnm.Graph(ctx, m, M)
        p = g.pack(V, dt)
        info = p.info()
        B = vnm.gen_normal(23, (m.ncols, n))
        tdt = torch.float32 if dt == vnm.DT_F32 else torch.bfloat16
        Bd = torch.from_numpy(B).to(tdt).cuda()
        C_ref = torch.full((m.nrows, n), float("nan"), device="cuda")
        C = torch.full((m.nrows, n), float("nan"), device="cuda")
        torch.cuda.synchronize()  # inputs made on torch's stream are complete
        p.spmm(Bd.data_ptr(), n, n, dt, C_ref.data_ptr(), n)  # tunes, loads every kernel
        ref64 = oracle.spmm_vnm_f64(_om(m), V, M, Bd.float().cpu().numpy())
        prepared.append((g, p, info, Bd, C_ref, C, n, dt, tol, ref64))
    assert prepared[0][2].slots > 0  # hub rows cut into chunks: the slots are used
    torch.cuda.synchronize()
    for k, (g, p, info, Bd, C_ref, C, n, dt, tol, ref64) in enumerate(prepared):
        print(f"call {k}: nrows {info.nrows} items {info.items} slots {info.slots} n {n} dtype {dt}", flush=True)
        free0 = torch.cuda.mem_get_info()[0]
        torch.cuda.synchronize()  # inputs made on torch's stream are complete
        p.spmm_noalloc(Bd.data_ptr(), n, n, dt, C.data_ptr(), n)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        got = C.cpu().numpy()
        assert np.array_equal(got.view(np.int32), C_ref.cpu().numpy().view(np.int32)), k
        fro, mx = _errs(got, ref64)
        assert fro < tol and mx < tol, (k, fro, mx)
    # more slots x columns than the static array holds: refused before any launch
    g, p, info = prepared[0][0], prepared[0][1], prepared[0][2]
    n_big = ((8 << 20) // info.slots + 64) // 64 * 64
    assert info.slots * n_big > (8 << 20)
    Bbig = torch.empty((g.ncols, n_big), dtype=torch.float32, device="cuda")
    Cbig = torch.empty((g.nrows, n_big), dtype=torch.float32, device="cuda")
    with pytest.raises(vnm.VnmError) as e:
        p.spmm_noalloc(Bbig.data_ptr(), n_big, n_big, vnm.DT_F32, Cbig.data_ptr(), n_big)
    assert e.value.code == vnm.E_ARG
    for g, p, *_ in prepared:
        p.close()
        g.close()
    ctx.close()
