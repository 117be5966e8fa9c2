"""The benchmark workloads at their full sizes.

C3 (64 x 2^22, skewed / duplicate-heavy, k = 16384) is checked against the
oracle bit-exactly, on the sampled and the exact path.  C4 (2048 uneven rows,
3.39 G keys, fp32, bf16 and f16) is checked against the oracle over the whole
batch (13.6 GB of fp32 values copied to the host; the multithreaded oracle
takes about a second on 16 cores)."""

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
def test_c4_full_size_against_oracle(dt):
    """C4 at its full size (2048 uneven rows, 3.39 G keys) in every dtype:
    values and indices bit-exact against the oracle, which runs over the
    whole batch on the host (all threads of this process's CPU share)."""
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W
    from helpers import DT_CODE, from_device_bits
    from oracle import oracle as orc
    w = W.c4(dt)
    vals = W.materialize(w)
    off = w.offsets()
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, total_len=w.total, max_k_=w.max_k)
    st = rk.status(ws)
    assert st[0] == 0, st
    got_b = from_device_bits(ov, dt)
    got_i = oi.cpu().numpy()
    del ov, oi, ws
    bits = from_device_bits(vals, dt)[: w.total]
    del vals
    torch.cuda.empty_cache()
    eb, ei = orc.topk(bits, off, w.ks, dt=DT_CODE[dt], flags=orc.LARGEST | orc.SORTED,
                      nthreads=orc.usable_threads())
    np.testing.assert_array_equal(got_i, ei)
    np.testing.assert_array_equal(got_b, eb)
