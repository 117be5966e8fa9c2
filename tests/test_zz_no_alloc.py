"""The allocation-free entry (rk_topk_batched_noalloc) against the batched
entry with its own workspace, through the C-ABI: two calls in a row on the
shared device scratch, the second smaller (and on another path), each
bit-exact against rk_topk_batched and the oracle; an oversize call is
refused before any device work."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import DT_CODE, from_device_bits, gpu_topk, to_device_bits
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LARGEST, SORTED, ASCEND, EXACT = 1, 2, 4, 8


def _noalloc(dt, bits, off, ks, flags):
    import torch

    import radik_amd as rk
    lens = off[1:] - off[:-1]
    oo = np.zeros(len(ks) + 1, np.int64)
    oo[1:] = np.cumsum(np.minimum(lens, ks.astype(np.int64)))
    n_out = int(oo[-1])
    vals = to_device_bits(bits, dt)
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(ks).cuda()
    d_oo = torch.from_numpy(oo).cuda()
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dt]
    ov = torch.zeros(max(n_out, 1), dtype=tdt, device="cuda")
    oi = torch.full((max(n_out, 1),), -7, dtype=torch.int32, device="cuda")
    total, mk = int(off[-1]), int(ks.max())
    need = rk.workspace_size(DT_CODE[dt], len(ks), total, mk, flags)
    assert need <= rk.RK_NOALLOC_SCRATCH_BYTES, need
    rc = rk.lib().rk_topk_batched_noalloc(DT_CODE[dt], vals.data_ptr(), d_off.data_ptr(),
                                          d_k.data_ptr(), len(ks), total, mk, flags, None,
                                          ov.data_ptr(), oi.data_ptr(), d_oo.data_ptr(),
                                          torch.cuda.current_stream().cuda_stream)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return from_device_bits(ov, dt)[:n_out], oi.cpu().numpy()[:n_out]


def _rows(seed, nrows, lo, hi, kmax):
    rng = np.random.default_rng(seed)
    lens = np.floor(2.0 ** rng.uniform(np.log2(lo), np.log2(hi), nrows)).astype(np.int64)
    off = np.zeros(nrows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    ks = rng.integers(1, kmax, nrows).astype(np.int32)
    return rng, off, ks


def test_noalloc_matches_batched_two_calls_second_smaller():
    # call 1: 96 uneven rows up to 2^19 (sampled filter path), fp32 largest sorted
    rng, off, ks = _rows(5, 96, 1000, 1 << 19, 4096)
    bits = rng.standard_normal(int(off[-1])).astype(np.float32).view(np.uint32)
    f1 = LARGEST | SORTED
    nb, ni = _noalloc("f32", bits, off, ks, f1)
    bb, bi, st = gpu_topk("f32", bits, off, ks, f1)
    np.testing.assert_array_equal(ni, bi)
    np.testing.assert_array_equal(nb, bb)
    ob, oi = orc.topk(bits, off, ks, dt=0, flags=f1, nthreads=8)
    np.testing.assert_array_equal(ni, oi)
    np.testing.assert_array_equal(nb, ob)
    # call 2: smaller (7 rows, bf16, smallest-k unsorted, exact passes), on the
    # scratch the first call left behind
    rng, off2, ks2 = _rows(6, 7, 100, 1 << 16, 2000)
    bits2 = rng.integers(0, 1 << 16, int(off2[-1])).astype(np.uint16)
    bits2[::5] = bits2[0]  # ties
    f2 = ASCEND | EXACT
    nb2, ni2 = _noalloc("bf16", bits2, off2, ks2, f2)
    bb2, bi2, _ = gpu_topk("bf16", bits2, off2, ks2, f2)
    np.testing.assert_array_equal(ni2, bi2)
    np.testing.assert_array_equal(nb2, bb2)
    ob2, oi2 = orc.topk(bits2, off2, ks2, dt=2, flags=f2 & 7, nthreads=8)
    np.testing.assert_array_equal(ni2, oi2)
    np.testing.assert_array_equal(nb2, ob2)


def test_noalloc_refuses_a_call_beyond_its_scratch():
    """A call whose workspace exceeds RK_NOALLOC_SCRATCH_BYTES returns
    RK_E_SCRATCH and leaves the outputs untouched."""
    import torch

    import radik_amd as rk
    rows, n = 4096, 1 << 16
    total = rows * n
    assert rk.workspace_size(rk.RK_F32, rows, total, 4096, 3) > rk.RK_NOALLOC_SCRATCH_BYTES
    d_off = torch.arange(rows + 1, dtype=torch.int64, device="cuda") * n
    d_k = torch.full((rows,), 4096, dtype=torch.int32, device="cuda")
    d_oo = torch.arange(rows + 1, dtype=torch.int64, device="cuda") * 4096
    vals = torch.zeros(1, device="cuda")  # never read: refused on the host
    ov = torch.full((8,), 3.0, device="cuda")
    oi = torch.full((8,), 5, dtype=torch.int32, device="cuda")
    rc = rk.lib().rk_topk_batched_noalloc(rk.RK_F32, vals.data_ptr(), d_off.data_ptr(),
                                          d_k.data_ptr(), rows, total, 4096, 3, None,
                                          ov.data_ptr(), oi.data_ptr(), d_oo.data_ptr(), None)
    assert rc == rk.RK_E_SCRATCH
    torch.cuda.synchronize()
    assert torch.equal(ov, torch.full((8,), 3.0, device="cuda"))
    assert torch.equal(oi, torch.full((8,), 5, dtype=torch.int32, device="cuda"))
