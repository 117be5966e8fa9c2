"""Unsorted output with k above the LDS sort limit (rk_max_k()): medians and
other quantile-sized selections (k up to len / 2 and beyond) through the
sampled bracket of the k-th key (or, on the exact path and when the bracket
misses, the digit passes) and the ordered emit, bit-exact against the
oracle's position-order output, for every dtype, both directions, ties,
idx_in, extra tail slots and rows copied whole; sorted calls with such k are
refused."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import DT_CODE, from_device_bits, gpu_topk
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LARGEST, SORTED, ASCEND, EXACT = 1, 2, 4, 8


def _bits(dt, kind, n, seed, a=0.0, b=1.0, dup=0.0):
    import torch

    import radik_amd as rk
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dt]
    t = torch.empty(max(n, 1), dtype=tdt, device="cuda")
    rk.gen_fill(t[:n], kind, seed, 0, a=a, b=b, dup_frac=dup, row_len=n)
    torch.cuda.synchronize()
    return from_device_bits(t, dt)[:n]


def _check(dt, bits, off, ks, flags, out_off=None, idx_in=None):
    gb, gi, st = gpu_topk(dt, bits, off, ks, flags, out_off=out_off, idx_in=idx_in)
    ob, oi = orc.topk(bits, off, ks, dt=DT_CODE[dt], flags=flags & 7, idx_in=idx_in,
                      out_off=out_off, nthreads=8)
    if not np.array_equal(gi, oi) or not np.array_equal(gb, ob):
        bad = np.nonzero((gi != oi) | (gb != ob))[0]
        raise AssertionError(f"{dt} flags={flags} mismatch at {bad[:10]} of {len(gi)}; "
                             f"gpu {gi[bad[:5]]} oracle {oi[bad[:5]]} status {st}")
    return st


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("largest", [True, False])
def test_median_and_large_k_unsorted(dt, largest):
    import radik_amd as rk
    lens = np.array([1 << 20, 300_001, 65_536, 40_000, 1 << 18], np.int64)
    ks = np.array([1 << 19, 150_000, 20_000, 16_385, 200_000], np.int32)  # medians and more
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    bits = _bits(dt, rk.GEN_NORMAL, int(off[-1]), 91)
    for exact in (False, True):
        st = _check(dt, bits, off, ks, (LARGEST if largest else 0) | (EXACT if exact else 0))
        assert st[2] >= 5  # every row took the ordered emit


def test_large_k_ties_idx_in_tail_slots_and_copy_all():
    import radik_amd as rk
    # near-constant rows with duplicates, an all-equal row, a row copied whole
    # (len <= k > rk_max_k()), mixed with ordinary rows in the same call
    lens = np.array([1 << 19, 100_000, 50_000, 30_000, 5_000], np.int64)
    ks = np.array([300_000, 70_000, 60_000, 17, 4_000], np.int32)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    n = int(off[-1])
    bits = _bits("f32", rk.GEN_NEARCONST, n, 92, dup=0.3)
    bits[off[1]:off[2]] = np.float32(0.25).view(np.uint32)  # all equal
    rng = np.random.default_rng(93)
    idx_in = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    m = np.minimum(lens, ks.astype(np.int64))
    out_off = np.zeros(len(lens) + 1, np.int64)
    out_off[1:] = np.cumsum(m + np.array([3, 0, 11, 0, 2]))  # extra slots: tail fill
    for largest in (True, False):
        _check("f32", bits, off, ks, LARGEST if largest else 0, out_off=out_off)
        _check("f32", bits, off, ks, LARGEST if largest else 0, out_off=out_off, idx_in=idx_in)


def test_compat_and_torch_op_large_k_unsorted():
    import torch

    import radik_amd as rk
    import radik_amd.torch_ops  # noqa: F401
    task_len = [70_000, 30_000, 100_000]
    K = 40_000
    off = np.zeros(len(task_len) + 1, np.int64)
    off[1:] = np.cumsum(task_len)
    bits = _bits("f32", rk.GEN_UNIFORM, int(off[-1]), 94)
    vals = torch.from_numpy(bits.view(np.int32).copy()).cuda().view(torch.float32)
    vo, io = rk.topk_radix_select(vals, K, task_len, largest=True, need_sorting=False)
    ob, oi = orc.topk(bits, off, None, K=K, flags=orc.LARGEST)
    np.testing.assert_array_equal(io.cpu().numpy().ravel(), oi)
    np.testing.assert_array_equal(vo.cpu().numpy().view(np.uint32).ravel(), ob)
    # torch op: unsorted top-k with k = 40000 of 100000 distinct values
    x = (torch.randperm(100_000, generator=torch.Generator().manual_seed(5)).float() - 5e4).cuda()
    v, i = torch.ops.radik_amd.topk(x, K, -1, True, False)
    ri = torch.topk(x, K).indices
    assert torch.equal(i, torch.sort(ri).values)
    assert torch.equal(v, x[i])
    with pytest.raises(ValueError):
        torch.ops.radik_amd.topk(x, K, -1, True, True)  # sorted: beyond the LDS sort


def test_sorted_large_k_refused():
    import torch

    import radik_amd as rk
    n, k = 100_000, rk.max_k() + 1
    vals = torch.zeros(n, device="cuda")
    off = torch.tensor([0, n], dtype=torch.int64, device="cuda")
    ks = torch.tensor([k], dtype=torch.int32, device="cuda")
    with pytest.raises(rk.RadikError) as e:
        rk.topk_batched(vals, off, ks, sorted=True, total_len=n, max_k_=k)
    assert e.value.code == rk.RK_E_UNSUPPORTED
    v, i, _, ws = rk.topk_batched(vals, off, ks, sorted=False, total_len=n, max_k_=k)
    assert rk.status(ws)[0] == 0
    torch.cuda.synchronize()
    assert torch.equal(i[:k].cpu(), torch.arange(k, dtype=torch.int32))  # all ties: positions


@pytest.mark.parametrize("dt", ["f32", "bf16"])
@pytest.mark.parametrize("largest", [True, False])
def test_bracket_missed_by_the_sample_falls_back(dt, largest):
    """Rows whose sampled runs are unrepresentative: values planted on exactly
    the positions the sample reads put the k-th key outside the sampled
    bracket (above it in one row, below it in another), beside a row the
    bracket settles.  The missed rows take the digit passes and every row
    still equals the oracle."""
    import radik_amd as rk
    import radik_amd.workloads as W
    n = 1 << 18
    esz = 4 if dt == "f32" else 2
    pack = 16 // esz
    lens = np.full(3, n, np.int64)
    off = np.zeros(4, np.int64)
    off[1:] = np.cumsum(lens)
    ks = np.array([n // 2, n // 4, n // 3], np.int32)
    bits = _bits(dt, rk.GEN_NORMAL, int(off[-1]), 57)
    big = 100.0 if largest else -100.0
    conv = (lambda v: np.float32(v).view(np.uint32)) if dt == "f32" else \
        (lambda v: np.uint16(np.float32(v).view(np.uint32) >> 16))
    for r, v in ((0, big), (1, -big)):
        a0 = ((-int(off[r]) * esz) % 16) // esz  # the row's first 16-byte aligned element
        pos = W.sample_positions(n, 1024, 16, pack, a0, full=True)
        bits[int(off[r]) + pos] = conv(v)
    st = _check(dt, bits, off, ks, LARGEST if largest else 0)
    assert st[2] == 3  # every row took the ordered emit
