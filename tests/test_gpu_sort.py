"""GPU parity of the library's stable radix sort (vnm_sort_pairs).

The order must be numpy's stable argsort of the keys' low end_bit bits, bit
for bit: keys move whole (bits above end_bit ride along), equal sort keys
keep their input order (checked through the values, which start as the
input positions).  Sizes around the 4,096-key tile and 64-key round edges,
low-entropy keys (long runs of one digit), digit widths that do not divide
into 8-bit passes, keys only and key-value pairs.
"""
import numpy as np
import pytest

import vnm

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    c = vnm.Context(0)
    yield c
    c.close()


def _keys(rng, n, kind):
    if kind == "wide":
        return rng.integers(0, 2 ** 63, n, dtype=np.int64).view(np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, n).astype(np.uint64)
    if kind == "few":  # a handful of distinct values: long equal runs
        return rng.choice(np.array([0, 1, 255, 256, 2 ** 40 + 7, 2 ** 63 + 5], np.uint64), n)
    return (rng.integers(0, 1 << 20, n).astype(np.uint64) << np.uint64(32)) | \
        rng.integers(0, 1000, n).astype(np.uint64)  # (row, col) keys as the packer builds them


# one tile (4,096 keys) and below: one launch in LDS; 2-4 tiles: one
# workgroup over global buffers; more: three launches per pass
@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 4096, 4097, 8192, 12289, 16384, 16385, 70001, 1_000_003])
def test_sort_pairs_matches_stable_argsort(ctx, n):
    rng = np.random.default_rng(n)
    for kind in ("wide", "few", "rowcol"):
        keys = _keys(rng, n, kind)
        for end_bit in (1, 7, 8, 13, 40, 52, 64):
            mask = np.uint64((1 << end_bit) - 1) if end_bit < 64 else np.uint64(2 ** 64 - 1)
            order = np.argsort(keys & mask, kind="stable")
            kin = torch.from_numpy(keys.view(np.int64)).cuda()
            vin = torch.arange(n, dtype=torch.int32, device="cuda")
            kout = torch.empty_like(kin)
            vout = torch.empty_like(vin)
            torch.cuda.synchronize()  # inputs made on torch's stream are complete
            ctx.sort_pairs(kin.data_ptr(), kout.data_ptr(), n, end_bit, vin.data_ptr(), vout.data_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(kout.cpu().numpy().view(np.uint64), keys[order]), (kind, end_bit)
            assert np.array_equal(vout.cpu().numpy(), order.astype(np.int32)), (kind, end_bit)
            assert np.array_equal(kin.cpu().numpy().view(np.uint64), keys)  # input untouched
            if end_bit in (13, 64):
                k2 = torch.empty_like(kin)
                torch.cuda.synchronize()  # inputs made on torch's stream are complete
                ctx.sort_pairs(kin.data_ptr(), k2.data_ptr(), n, end_bit)
                torch.cuda.synchronize()
                assert np.array_equal(k2.cpu().numpy().view(np.uint64), keys[order]), (kind, end_bit)


def test_sort_rejects_bad_arguments(ctx):
    k = torch.zeros(8, dtype=torch.int64, device="cuda")
    for bad in (dict(end_bit=0), dict(end_bit=65)):
        with pytest.raises(vnm.VnmError):
            ctx.sort_pairs(k.data_ptr(), k.data_ptr() + 64, 8, **bad)
    with pytest.raises(vnm.VnmError):
        ctx.sort_pairs(k.data_ptr(), k.data_ptr(), 8, 64)  # in place
