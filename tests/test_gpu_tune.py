"""Tunables and the autotuner: every setting the autotuner (or a caller, via
rk_set_tune) can choose returns the same bit-exact results; the autotuner
caches a winner per shape bucket, which later calls use and which survives a
save / clear / load of the cache file."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cus():
    import torch
    return torch.cuda.get_device_properties(0).multi_processor_count


def _case(name):
    """(workload, device values, device offsets, device ks, expected bits, idx)."""
    import torch

    import radik_amd.workloads as W
    from helpers import from_device_bits
    from oracle import oracle as orc
    w = {"c4": lambda: W.c4("f32", scale=64, rows=256),
         "c3": lambda: W.c3(scale=64, rows=8)}[name]()
    vals = W.materialize(w)
    d_off = torch.from_numpy(w.offsets()).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    flags = (orc.LARGEST if w.largest else 0) | orc.SORTED
    eb, ei = orc.topk(from_device_bits(vals, "f32")[: w.total], w.offsets(), w.ks, flags=flags,
                      nthreads=8)
    return w, vals, d_off, d_k, eb, ei


def _run_check(w, vals, d_off, d_k, eb, ei, exact):
    import radik_amd as rk
    from helpers import from_device_bits
    ov, oi, _, ws = rk.topk_batched(vals, d_off, d_k, largest=w.largest, exact=exact,
                                    total_len=w.total, max_k_=w.max_k)
    st = rk.status(ws)
    assert st[0] == 0, st
    np.testing.assert_array_equal(oi.cpu().numpy()[: len(ei)], ei)
    np.testing.assert_array_equal(from_device_bits(ov, "f32")[: len(eb)], eb)
    return st


def _tunes():
    c = _cus()
    return [
        dict(tile=4096, grid=4 * c, unroll=2, sample_groups=512, group_elems=16, hist_bits=8),
        dict(tile=65536, grid=16 * c, unroll=8, sample_groups=2048, group_elems=16, hist_bits=9),
        dict(tile=16384, grid=7, unroll=4, sample_groups=64, group_elems=256, hist_bits=10),
        # smallest settings: one persistent workgroup, minimum tile and sample
        dict(tile=1024, grid=1, unroll=4, sample_groups=1, group_elems=8, hist_bits=11),
    ]


@pytest.mark.parametrize("name", ["c4", "c3"])
def test_every_tunable_setting_is_bit_exact(name):
    import radik_amd as rk
    case = _case(name)
    old = rk.get_tune()
    rk.tune_cache_clear()
    try:
        for t in _tunes():
            rk.set_tune(**t)
            assert rk.get_tune()["hist_bits"] == t["hist_bits"]
            for exact in (False, True):
                _run_check(*case, exact=exact)
    finally:
        rk.set_tune(**old)


def test_autotuner_caches_per_bucket_and_roundtrips(tmp_path):
    import torch

    import radik_amd as rk
    w, vals, d_off, d_k, eb, ei = _case("c4")
    oo = torch.from_numpy(w.out_offsets()).cuda()
    n = int(w.out_offsets()[-1])
    ov = torch.empty(n, dtype=torch.float32, device="cuda")
    oi = torch.empty(n, dtype=torch.int32, device="cuda")
    rk.tune_cache_clear()
    default = rk.get_tune()
    for exact in (False, True):
        grids = {m * _cus() for m in ((1, 2, 4) if exact else (2, 3, 4, 8))}
        ws = torch.empty(rk.workspace_size(rk.RK_F32, w.rows, w.total, w.max_k,
                                           rk.RK_LARGEST | rk.RK_SORTED
                                           | (rk.RK_EXACT if exact else 0)),
                         dtype=torch.uint8, device="cuda")
        best, ms = rk.autotune_batched(vals, d_off, d_k, ov, oi, oo, ws, total_len=w.total,
                                       max_k_=w.max_k, exact=exact, reps=1)
        assert ms > 0
        assert best["tile"] in (4096, 16384, 65536) and best["grid"] in grids
        assert best["unroll"] in (2, 4)
        if exact:
            assert best["hist_bits"] in (8, 11)
        else:
            assert best["sample_groups"] in (512, 2048)
        assert rk.get_tune() == default  # the defaults are restored
        _run_check(w, vals, d_off, d_k, eb, ei, exact)  # runs with the cached winner
    assert rk.tune_cache_size() == 2  # one bucket per path
    path = str(tmp_path / "radik.tune")
    rk.tune_cache_save(path)
    rk.tune_cache_clear()
    assert rk.tune_cache_size() == 0
    rk.tune_cache_load(path)
    assert rk.tune_cache_size() == 2
    lines = [l for l in open(path) if not l.startswith("#")]
    assert len(lines) == 2 and all(len(l.split()) == 11 for l in lines)
    for exact in (False, True):
        _run_check(w, vals, d_off, d_k, eb, ei, exact)
    rk.tune_cache_clear()
