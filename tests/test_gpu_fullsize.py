"""The benchmark workloads at their full sizes.

C3 (64 x 2^22, skewed / duplicate-heavy, k = 16384) is checked against the
oracle bit-exactly.  C4 (2048 uneven rows, 3.39 G keys, fp32, bf16 and f16) is too
large for the oracle in a test, so every row is checked on the device with
size-independent properties: the selected values equal torch.topk's values
(the multiset of the k largest), every index points at its value, and the
ties at the k-th value are the lowest positions holding it (the rule the
oracle and the reference's semantics fix)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c3_full_size_against_oracle():
    import radik_amd as rk
    import radik_amd.workloads as W
    from helpers import from_device_bits
    from oracle import oracle as orc
    w = W.c3()
    vals = W.materialize(w)
    off = w.offsets()
    bits = from_device_bits(vals, "f32")[: w.total]
    import torch
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    eb, ei = orc.topk(bits, off, w.ks, flags=orc.LARGEST | orc.SORTED, nthreads=8)
    for exact in (False, True):
        ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, exact=exact, total_len=w.total,
                                        max_k_=w.max_k)
        assert rk.status(ws)[0] == 0
        np.testing.assert_array_equal(oi.cpu().numpy(), ei)
        np.testing.assert_array_equal(from_device_bits(ov, "f32"), eb)


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
def test_c4_full_size_properties(dt):
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W
    w = W.c4(dt)
    vals = W.materialize(w)
    off = w.offsets()
    oo = w.out_offsets()
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, total_len=w.total, max_k_=w.max_k)
    st = rk.status(ws)
    assert st[0] == 0, st
    oi64 = oi.long()
    for r in range(w.rows):
        lo, hi = int(off[r]), int(off[r + 1])
        a, b = int(oo[r]), int(oo[r + 1])
        row = vals[lo:hi].float()
        k = b - a
        got_v = ov[a:b].float()
        got_i = oi64[a:b]
        ref = torch.topk(row, k, sorted=True).values
        assert torch.equal(got_v, ref), f"row {r}: values differ"
        assert torch.equal(row[got_i], got_v), f"row {r}: index does not point at its value"
        kth = ref[-1]
        n_gt = int((row > kth).sum())
        # the k - n_gt ties chosen are the lowest positions equal to kth,
        # output after the strictly larger values in position order
        eq_pos = torch.nonzero(row == kth).flatten()[: k - n_gt]
        assert torch.equal(got_i[n_gt:], eq_pos), f"row {r}: ties not the lowest positions"
        assert torch.equal(torch.sort(got_i[:n_gt]).values, torch.nonzero(row > kth).flatten())
