"""Inputs built against the sampled path's sample (``W.adversarial``): the
best keys hidden between the sampled runs (the candidates overflow) or placed
only on them (too few keys reach the threshold).  Every row falls back to the
exact passes, and the output must still equal the oracle's, values and
indices, in both directions and in fp32 and bf16."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", ["f32", "bf16"])
@pytest.mark.parametrize("largest", [True, False])
@pytest.mark.parametrize("kind", ["hidden", "decoy"])
def test_sampler_defeating_rows_fall_back_exactly(kind, largest, dt):
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W
    from helpers import DT_CODE, from_device_bits
    from oracle import oracle as orc
    rk.set_tune(tile=0, grid=0, sample_groups=0, group_elems=0, hist_bits=0, unroll=0)
    w = W.adversarial(kind, rows=8, n=1 << 20, k=1024, dtype=dt, largest=largest)
    vals = W.materialize(w)
    off = w.offsets()
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    bits = from_device_bits(vals, dt)[: w.total]
    fl = (orc.LARGEST if largest else orc.ASCEND) | orc.SORTED  # best first, as topk_batched
    eb, ei = orc.topk(bits, off, w.ks, dt=DT_CODE[dt], flags=fl, nthreads=orc.usable_threads())
    ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, largest=largest, total_len=w.total,
                                    max_k_=w.max_k)
    st = rk.status(ws)
    assert st[0] == 0, st
    assert st[1] == w.rows, f"every row should take the exact passes: {st}"
    np.testing.assert_array_equal(oi.cpu().numpy(), ei)
    np.testing.assert_array_equal(from_device_bits(ov, dt), eb)
    # the same rows on the exact path
    ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, largest=largest, exact=True,
                                    total_len=w.total, max_k_=w.max_k)
    assert rk.status(ws)[0] == 0
    np.testing.assert_array_equal(oi.cpu().numpy(), ei)
    np.testing.assert_array_equal(from_device_bits(ov, dt), eb)
