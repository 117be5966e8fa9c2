"""Serving loop (rk_server_*): every request's result is the same alone and
in the mix (bit-exact, and equal to the oracle), requests are coalesced into
batches, the memory budget bounds what is in flight, and the plain baseline
returns the same results."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_K = 4096


def _requests(n=96, seed=21, const_rows=2):
    """Mixed-size requests: lengths log-uniform in [2^8, 2^18], k log-uniform
    in [1, 2^12] (so some k exceed the row), the four mixture families, plus
    constant rows (all ties: they leave the sampled path for the exact passes
    and the ordered emit)."""
    from radik_amd.workloads import uniform01
    lens = np.floor(2.0 ** (8 + 10 * uniform01(seed, 1, n))).astype(np.int64)
    ks = np.maximum(np.floor(2.0 ** (12 * uniform01(seed, 2, n))).astype(np.int32), 1)
    kinds = np.floor(4 * uniform01(seed, 3, n)).astype(np.int32)
    lens[:const_rows] = 1 << 16
    ks[:const_rows] = [100, 3000][:const_rows]
    return lens, ks, kinds


def _arena(lens, kinds, const_rows):
    import radik_amd as rk
    import radik_amd.workloads as W
    vals, off = W.materialize_requests(lens, kinds, seed=21)
    for i in range(const_rows):
        lo, hi = int(off[i]), int(off[i + 1])
        rk.gen_fill(vals[lo:hi], rk.GEN_CONST, 21, i, a=0.75, b=0.75, row_len=hi - lo)
    return vals, off


def _outputs(ks):
    import torch
    oo = np.zeros(len(ks) + 1, np.int64)
    oo[1:] = np.cumsum(ks.astype(np.int64))
    ov = torch.full((int(oo[-1]),), 7.0, dtype=torch.float32, device="cuda")
    oi = torch.full((int(oo[-1]),), -7, dtype=torch.int32, device="cuda")
    return oo, ov, oi


def _ptrs(t, off, esize):
    return [t.data_ptr() + int(o) * esize for o in off[:-1]]


def _expected(vals, off, ks):
    from helpers import from_device_bits
    from oracle import oracle as orc
    bits = from_device_bits(vals, "f32")[: int(off[-1])]
    oo = np.zeros(len(ks) + 1, np.int64)
    oo[1:] = np.cumsum(ks.astype(np.int64))
    return orc.topk(bits, off, ks, flags=orc.LARGEST | orc.SORTED, out_off=oo, nthreads=8)


def _bits(t):
    import torch
    return t.view(torch.int32).cpu().numpy().view(np.uint32)


def test_alone_and_in_the_mix_are_identical():
    import torch

    import radik_amd as rk
    lens, ks, kinds = _requests()
    vals, off = _arena(lens, kinds, 2)
    torch.cuda.synchronize()
    eb, ei = _expected(vals, off, ks)
    n = len(lens)

    # alone: each request submitted and waited for before the next
    oo, av, ai = _outputs(ks)
    with rk.Server(max_k=MAX_K, max_batch_rows=256, max_batch_keys=1 << 24, streams=4) as srv:
        for i in range(n):
            lo, hi = int(off[i]), int(off[i + 1])
            t = srv.submit(vals[lo:hi], int(ks[i]), av[int(oo[i]):], ai[int(oo[i]):])
            srv.wait(t)
        st = srv.stats()
        assert st["completed"] == n and st["max_batch_rows"] == 1
        assert st["fallback_batches"] >= 2  # the constant rows took the exact passes
    np.testing.assert_array_equal(ai.cpu().numpy(), ei)
    np.testing.assert_array_equal(_bits(av), eb)

    # in the mix: a burst (all at t=0), then Poisson arrivals at a high rate,
    # so requests arrive while earlier batches are on the GPU
    for arrivals in (np.zeros(n), np.cumsum(np.full(n, 20e-6))):
        _, mv, mi = _outputs(ks)
        with rk.Server(max_k=MAX_K, max_batch_rows=256, max_batch_keys=1 << 24,
                       streams=4) as srv:
            lat, makespan = srv.replay(_ptrs(vals, off, 4), lens, ks, _ptrs(mv, oo, 4),
                                       _ptrs(mi, oo, 4), arrivals)
            st = srv.stats()
        assert st["completed"] == n and st["failed"] == 0
        assert np.all(lat >= 0) and makespan > 0
        if arrivals[-1] == 0:
            assert st["mean_batch_rows"] > 1.0, st  # coalesced
        np.testing.assert_array_equal(mi.cpu().numpy(), ai.cpu().numpy())
        np.testing.assert_array_equal(_bits(mv), _bits(av))


def test_plain_baseline_same_results():
    import torch

    import radik_amd as rk
    lens, ks, kinds = _requests(n=40, seed=22, const_rows=1)
    vals, off = _arena(lens, kinds, 1)
    torch.cuda.synchronize()
    eb, ei = _expected(vals, off, ks)
    oo, pv, pi = _outputs(ks)
    with rk.Server(max_k=MAX_K, max_batch_keys=1 << 24, plain=True) as srv:
        lat, _ = srv.replay(_ptrs(vals, off, 4), lens, ks, _ptrs(pv, oo, 4), _ptrs(pi, oo, 4),
                            np.zeros(len(lens)))
        st = srv.stats()
    assert st["max_batch_rows"] == 1 and st["batches"] == len(lens) and st["failed"] == 0
    np.testing.assert_array_equal(pi.cpu().numpy(), ei)
    np.testing.assert_array_equal(_bits(pv), eb)


def test_admission_budget_bounds_inflight_and_rejects_oversize():
    import torch

    import radik_amd as rk
    n = 24
    lens = np.full(n, 1 << 16, np.int64)
    ks = np.full(n, 64, np.int32)
    kinds = np.arange(n, dtype=np.int32) % 4
    lens[5] = 1 << 20  # larger than the whole request budget below
    vals, off = _arena(lens, kinds, 0)
    torch.cuda.synchronize()
    eb, ei = _expected(vals, off, ks)
    oo, ov, oi = _outputs(ks)
    req_bytes = (1 << 16) * 4 + 64 * 8
    kw = dict(max_k=256, max_batch_rows=64, max_batch_keys=1 << 21, streams=2)
    with rk.Server(**kw) as probe:
        server_bytes = probe.stats()["server_bytes"]
    with rk.Server(mem_budget=server_bytes + 3 * req_bytes, **kw) as srv:
        tickets = []
        for i in range(n):
            lo, hi = int(off[i]), int(off[i + 1])
            tickets.append(srv.submit(vals[lo:hi], int(ks[i]), ov[int(oo[i]):],
                                      oi[int(oo[i]):]))
        codes = [srv.wait_code(t) for t in tickets]
        st = srv.stats()
    assert codes[5] == rk.RK_E_WORKSPACE
    assert all(c == 0 for i, c in enumerate(codes) if i != 5)
    assert st["rejected"] == 1 and st["completed"] == n - 1
    assert 0 < st["peak_inflight_bytes"] <= 3 * req_bytes
    keep = np.ones(int(oo[-1]), bool)
    keep[int(oo[5]):int(oo[6])] = False
    np.testing.assert_array_equal(oi.cpu().numpy()[keep], ei[keep])
    np.testing.assert_array_equal(_bits(ov)[keep], eb[keep])


def test_smallest_bf16_server():
    import torch

    import radik_amd as rk
    from helpers import from_device_bits
    from oracle import oracle as orc
    lens, ks, kinds = _requests(n=32, seed=23, const_rows=0)
    vals, off = _arena(lens, kinds, 0)
    vb = vals.to(torch.bfloat16)
    torch.cuda.synchronize()
    oo = np.zeros(len(ks) + 1, np.int64)
    oo[1:] = np.cumsum(ks.astype(np.int64))
    eb, ei = orc.topk(from_device_bits(vb, "bf16"), off, ks, dt=orc.DT_BF16,
                      flags=orc.SORTED | orc.ASCEND, out_off=oo, nthreads=8)
    ov = torch.zeros(int(oo[-1]), dtype=torch.bfloat16, device="cuda")
    oi = torch.zeros(int(oo[-1]), dtype=torch.int32, device="cuda")
    with rk.Server(rk.RK_BF16, largest=False, ascend=True, max_k=MAX_K,
                   max_batch_keys=1 << 24) as srv:
        srv.replay(_ptrs(vb, off, 2), lens, ks, _ptrs(ov, oo, 2), _ptrs(oi, oo, 4),
                   np.zeros(len(lens)))
        assert srv.stats()["failed"] == 0
    np.testing.assert_array_equal(oi.cpu().numpy(), ei)
    np.testing.assert_array_equal(from_device_bits(ov, "bf16"), eb)


def test_waited_requests_are_released():
    """The server releases a request once its outcome is reported: after
    submit + wait cycles nothing is retained, a second wait on a ticket is
    RK_E_ARG, and the statistics still cover every completed request."""
    import torch

    import radik_amd as rk
    n, rows = 1 << 16, 300
    vals = torch.randn(n, device="cuda")
    ov = torch.empty(rows * 16, device="cuda")
    oi = torch.empty(rows * 16, dtype=torch.int32, device="cuda")
    with rk.Server(max_k=16, max_batch_rows=8, max_batch_keys=1 << 20, streams=2) as srv:
        tickets = []
        for i in range(rows):
            tickets.append(srv.submit(vals[: 1000 + i * 100], 16, ov[16 * i:], oi[16 * i:]))
            if i % 3 == 2:
                for t in tickets:
                    srv.wait(t)
                tickets = []
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        assert st["completed"] == rows and st["retained"] == 0, st
        assert srv.wait_code(5) == rk.RK_E_ARG
    ref = torch.topk(vals[: 1000 + 299 * 100], 16).indices.to(torch.int32)
    assert torch.equal(oi[16 * 299:16 * 300], ref)
