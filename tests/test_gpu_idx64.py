"""64-bit indices (rk_topk_batched_i64 / rk_topk_compat_i64, the reference's
IdxType = int64_t, radik/RadixSelect/topk_radixselect.h:59-70): every path
of the pipeline writes int64 indices, gathered through int64 caller indices
beyond 2^31 or as row positions, with the int64 tail fill -1.  Expected
output: the oracle's values and positions, the positions mapped through the
same int64 array on the host."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import DT_CODE, from_device_bits, to_device_bits
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LARGEST, SORTED, ASCEND, EXACT = 1, 2, 4, 8


def _bits(dt, kind, n, seed, dup=0.0):
    import torch

    import radik_amd as rk
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dt]
    t = torch.empty(max(n, 1), dtype=tdt, device="cuda")
    rk.gen_fill(t[:n], kind, seed, 0, dup_frac=dup, row_len=n)
    torch.cuda.synchronize()
    return from_device_bits(t, dt)[:n]


def _expected(dt, bits, off, ks, flags, oo, idx64):
    eb, epos = orc.topk(bits, off, ks, dt=DT_CODE[dt], flags=flags & 7, out_off=oo, nthreads=8)
    ei = np.where(epos >= 0, idx64[np.maximum(epos, 0) + np.repeat(off[:-1], np.diff(oo))]
                  if idx64 is not None else epos.astype(np.int64), -1).astype(np.int64)
    return eb, ei


def _run_i64(dt, bits, off, ks, flags, oo, idx64):
    import torch

    import radik_amd as rk
    vals = to_device_bits(bits, dt)
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(ks).cuda()
    d_oo = torch.from_numpy(oo).cuda()
    ii = None if idx64 is None else torch.from_numpy(idx64).cuda()
    ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, largest=bool(flags & LARGEST),
                                    sorted=bool(flags & SORTED), ascend=bool(flags & ASCEND),
                                    exact=bool(flags & EXACT), idx_in=ii, out_off=d_oo,
                                    idx64=True)
    assert oi.dtype == torch.int64
    st = rk.status(ws)
    assert st[0] == 0, st
    n = int(oo[-1])
    return from_device_bits(ov, dt)[:n], oi.cpu().numpy()[:n], st


@pytest.mark.parametrize("dt", ["f32", "bf16"])
@pytest.mark.parametrize("flags", [LARGEST | SORTED, SORTED | ASCEND, LARGEST, LARGEST | EXACT,
                                   SORTED | ASCEND | EXACT])
@pytest.mark.parametrize("with_idx", [False, True])
def test_int64_indices_every_path(dt, flags, with_idx):
    # copy-all, LDS-direct, sampled, k > 4096 (chunked sort), all-equal rows
    # (ordered emit), near-constant duplicates, and extra tail slots
    lens = np.array([300, 9000, 500_000, 700_000, 1 << 18, 200_000, 100_000], np.int64)
    ks = np.array([1000, 700, 1000, 9000, 3000, 4096, 16], np.int32)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    n = int(off[-1])
    bits = _bits(dt, 1, n, 401 + flags)
    bits[off[4]:off[5]] = bits[off[4]]  # all equal: ties overflow the slots
    bits[off[5]:off[6]] = _bits(dt, 3, int(lens[5]), 402, dup=0.3)
    rng = np.random.default_rng(403)
    idx64 = rng.integers(-(1 << 62), 1 << 62, n, dtype=np.int64) if with_idx else None
    m = np.minimum(lens, ks.astype(np.int64))
    oo = np.zeros(len(lens) + 1, np.int64)
    oo[1:] = np.cumsum(m + np.array([5, 0, 0, 3, 0, 0, 2]))  # tail fill slots
    gb, gi, st = _run_i64(dt, bits, off, ks, flags, oo, idx64)
    eb, ei = _expected(dt, bits, off, ks, flags, oo, idx64)
    np.testing.assert_array_equal(gi, ei)
    np.testing.assert_array_equal(gb, eb)
    assert st[2] >= 1  # the all-equal row took the ordered emit


def test_int64_indices_unsorted_large_k():
    """k above the LDS select (unsorted): exact passes + ordered emit write
    int64 positions and map them through the int64 caller indices."""
    lens = np.array([1 << 20, 300_001], np.int64)
    ks = np.array([1 << 19, 20_000], np.int32)
    off = np.zeros(3, np.int64)
    off[1:] = np.cumsum(lens)
    bits = _bits("f32", 1, int(off[-1]), 404)
    idx64 = np.arange(int(off[-1]), dtype=np.int64) * 3 + (1 << 40)
    oo = np.zeros(3, np.int64)
    oo[1:] = np.cumsum(ks.astype(np.int64))
    for flags in (LARGEST, 0):
        gb, gi, _ = _run_i64("f32", bits, off, ks, flags, oo, idx64)
        eb, ei = _expected("f32", bits, off, ks, flags, oo, idx64)
        np.testing.assert_array_equal(gi, ei)
        np.testing.assert_array_equal(gb, eb)


def test_compat_int64():
    """rk_topk_compat_i64: [taskNum][K] int64 output with the -1 tail fill."""
    import torch

    import radik_amd as rk
    task_len = [5000, 70000, 12, 0, 40000]
    off = np.zeros(len(task_len) + 1, np.int64)
    off[1:] = np.cumsum(task_len)
    K = 2048
    bits = _bits("f32", 1, int(off[-1]), 405)
    vals = torch.from_numpy(bits.view(np.int32).copy()).cuda().view(torch.float32)
    idx64 = (np.arange(int(off[-1]), dtype=np.int64) << 33) + 7
    vo, io = rk.topk_radix_select(vals, K, task_len, largest=True, ascend=False,
                                  idxIn=torch.from_numpy(idx64).cuda())
    torch.cuda.synchronize()
    assert io.dtype == torch.int64
    oo = np.arange(len(task_len) + 1, dtype=np.int64) * K
    eb, epos = orc.topk(bits, off, None, K=K, flags=LARGEST | SORTED)
    base = np.repeat(off[:-1], K)
    ei = np.where(epos >= 0, idx64[np.maximum(epos, 0) + base], -1)
    np.testing.assert_array_equal(io.cpu().numpy().ravel(), ei)
    np.testing.assert_array_equal(vo.cpu().numpy().view(np.uint32).ravel(), eb)
