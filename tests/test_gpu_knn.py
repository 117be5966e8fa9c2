"""Fused k-nearest-neighbour select (rk_knn_topk): MFMA scores filtered
against a per-query threshold inside the score kernel, then selected exactly.

- The scores themselves (rk_knn_scores, the same arithmetic) against float64
  scores on the CPU, within a tolerance that scales with the magnitudes
  summed (fp32 accumulation of exact bf16/f16 products).
- rk_knn_topk bit for bit against the unfused path: those scores written to
  memory, then rk_topk_batched (largest for inner product, smallest for L2),
  sorted and unsorted, k from 1 to n, duplicate database rows (ties go to the
  lower index), sizes that are not multiples of the tiles.
- The second pass: a database whose sampled rows hold the best scores makes
  the first threshold admit fewer than k candidates.
- Candidate overflow (every score tied) is reported in the status word, and
  the failed queries' outputs are filled with NaN / -1.
- hipGraph capture: replays equal the eager result, also after the queries
  change in place.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(nq, n, d, dt, seed, dup=0):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    tdt = {"bf16": torch.bfloat16, "f16": torch.float16}[dt]
    q = torch.randn(nq, d, generator=g).to(tdt)
    x = torch.randn(n, d, generator=g).to(tdt)
    if dup:
        # copies of earlier rows: equal scores, to be ordered by index
        src = torch.randint(0, n // 2, (dup,), generator=g)
        dst = torch.randint(n // 2, n, (dup,), generator=g)
        x[dst] = x[src]
    return q.cuda(), x.cuda()


def _unfused(q, x, k, metric, sorted_):
    import torch

    import radik_amd as rk
    nq, n = q.shape[0], x.shape[0]
    s = rk.knn_scores(q, x, metric=metric)
    off = torch.arange(nq + 1, dtype=torch.int64, device="cuda") * n
    ks = torch.full((nq,), k, dtype=torch.int32, device="cuda")
    v, i, _, ws = rk.topk_batched(s.view(-1), off, ks, largest=(metric == rk.RK_KNN_IP),
                                  sorted=sorted_, total_len=nq * n, max_k_=k)
    assert rk.status(ws)[0] == 0
    return s, v.view(nq, k), i.view(nq, k)


@pytest.mark.parametrize("dt", ["bf16", "f16"])
@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("d", [16, 48, 128, 256])
def test_scores_against_float64(dt, metric, d):
    import radik_amd as rk
    q, x = _data(77, 1001, d, dt, 1 + d)
    s = rk.knn_scores(q, x, metric=metric).cpu().numpy().astype(np.float64)
    q64 = q.float().cpu().numpy().astype(np.float64)
    x64 = x.float().cpu().numpy().astype(np.float64)
    dot = q64 @ x64.T
    mag = np.abs(q64) @ np.abs(x64).T
    if metric == rk.RK_KNN_IP:
        ref, scale = dot, mag
    else:
        qn = (q64 * q64).sum(1)[:, None]
        xn = (x64 * x64).sum(1)[None, :]
        ref, scale = qn - 2 * dot + xn, qn + 2 * mag + xn
    err = np.abs(s - ref)
    assert np.all(err <= 5e-5 * scale + 1e-6), float((err / (scale + 1e-30)).max())


@pytest.mark.parametrize("dt", ["bf16", "f16"])
@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("sorted_", [True, False])
def test_fused_equals_unfused(dt, metric, sorted_):
    import radik_amd as rk
    q, x = _data(300, 50_001, 128, dt, 7, dup=2000)
    for k in (1, 10, 100, 1000):
        v, i, ws = rk.knn_topk(q, x, k, metric=metric, sorted=sorted_)
        st = rk.status(ws)
        assert st[0] == 0, st
        _, uv, ui = _unfused(q, x, k, metric, sorted_)
        np.testing.assert_array_equal(i.cpu().numpy(), ui.cpu().numpy())
        np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32),
                                      uv.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("n,k,d", [(33, 33, 16), (1000, 1, 64), (4096, 700, 256),
                                   (20_000, 16384, 32), (200_000, 64, 96)])
def test_shapes_and_k_extremes(n, k, d):
    import radik_amd as rk
    q, x = _data(129, n, d, "bf16", n + k)
    for metric in (rk.RK_KNN_IP, rk.RK_KNN_L2):
        v, i, ws = rk.knn_topk(q, x, k, metric=metric)
        assert rk.status(ws)[0] == 0
        _, uv, ui = _unfused(q, x, k, metric, True)
        np.testing.assert_array_equal(i.cpu().numpy(), ui.cpu().numpy())
        np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32),
                                      uv.cpu().numpy().view(np.uint32))


def test_second_pass_when_the_sample_misleads():
    import torch

    import radik_amd as rk
    # n = 100000 samples every 6th row (S = 16384); k = 1000 gives the
    # first threshold rank r = 218.  300 sampled rows score 100, the rest
    # about 0: only 300 < k rows reach the first threshold.
    n, k, d = 100_000, 1000, 64
    g = torch.Generator(device="cpu").manual_seed(3)
    x = (torch.randn(n, d, generator=g) * 0.01)
    x[torch.arange(300) * 6, 0] = 100.0
    q = torch.zeros(5, d)
    q[:, 0] = 1.0
    q[:, 1] = torch.arange(5).float() * 0.01
    q, x = q.to(torch.bfloat16).cuda(), x.to(torch.bfloat16).cuda()
    v, i, ws = rk.knn_topk(q, x, k)
    st = rk.status(ws)
    assert st[0] == 0 and st[1] == 5, st  # every query took the second pass
    _, uv, ui = _unfused(q, x, k, rk.RK_KNN_IP, True)
    np.testing.assert_array_equal(i.cpu().numpy(), ui.cpu().numpy())
    np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32),
                                  uv.cpu().numpy().view(np.uint32))


def test_all_tied_scores_report_overflow():
    import torch

    import radik_amd as rk
    n, k, d = 100_000, 10, 32
    x = torch.zeros(n, d, dtype=torch.bfloat16, device="cuda")
    q = torch.ones(3, d, dtype=torch.bfloat16, device="cuda")
    v, i, ws = rk.knn_topk(q, x, k)
    st = rk.status(ws)
    assert st[0] & 4 and st[1] == 3, st  # candidates outgrew the slots
    # the failed queries read as failed: value NaN, index -1
    assert torch.isnan(v).all() and bool((i == -1).all())


def test_arguments_rejected():
    import torch

    import radik_amd as rk
    q = torch.zeros(4, 64, dtype=torch.bfloat16, device="cuda")
    x = torch.zeros(100, 64, dtype=torch.bfloat16, device="cuda")
    with pytest.raises(rk.RadikError):
        rk.knn_topk(q, x, 101)  # k > n
    with pytest.raises(rk.RadikError):
        rk.knn_topk(q, x, 0)
    with pytest.raises(rk.RadikError):
        rk.knn_topk(q.float(), x.float(), 5)  # fp32 inputs are not MFMA inputs here
    with pytest.raises(rk.RadikError):
        rk.knn_topk(q[:, :40].contiguous(), x[:, :40].contiguous(), 5)  # d % 16


@pytest.mark.parametrize("metric", [0, 1])
def test_graph_replays(metric):
    """rk_knn_topk captured in a hipGraph: every replay (with eager calls on
    other buffers and synchronizations in between) equals the eager result,
    and a replay after the queries change in place equals the eager result
    on the new queries."""
    import torch

    import radik_amd as rk
    q, x = _data(200, 30_000, 64, "bf16", 21)
    k = 50
    ref_v, ref_i, _ = rk.knn_topk(q, x, k, metric=metric)
    torch.cuda.synchronize()
    ws = torch.empty(rk.knn_workspace_size(rk.RK_BF16, 200, 30_000, 64, k), dtype=torch.uint8,
                     device="cuda")
    out_v = torch.empty_like(ref_v)
    out_i = torch.empty_like(ref_i)
    s = torch.cuda.Stream()
    with rk.Graph.capture(s) as g:
        rk.knn_topk(q, x, k, metric=metric, workspace=ws, out_vals=out_v, out_idx=out_i, stream=s)
    for rep in range(3):
        out_v.fill_(0)
        out_i.fill_(-5)
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        assert rk.status(ws)[0] == 0
        assert torch.equal(out_i, ref_i), f"replay {rep}: indices differ"
        assert torch.equal(out_v.view(torch.int32), ref_v.view(torch.int32)), f"replay {rep}"
        ev, ei, _ = rk.knn_topk(q, x, k, metric=metric)  # eager, own buffers
        torch.cuda.synchronize()
        assert torch.equal(ei, ref_i)
    q.copy_(torch.flip(q, dims=[0]))
    new_v, new_i, _ = rk.knn_topk(q, x, k, metric=metric)
    torch.cuda.synchronize()
    g.launch(s)
    s.synchronize()
    assert torch.equal(out_i, new_i)
    assert torch.equal(out_v.view(torch.int32), new_v.view(torch.int32))
    g.close()


@pytest.mark.parametrize("nq,n,k", [(1, 50, 50), (2, 31, 7), (1, 16385, 1), (3, 100_000, 16384)])
def test_tiny_query_counts_and_k_equal_n(nq, n, k):
    """One or two queries, databases below one 32-vector tile, k = n, and the
    largest k: fused equals unfused for both metrics and both orders."""
    import radik_amd as rk
    q, x = _data(nq, n, 32, "f16", nq * 7 + n)
    for metric in (rk.RK_KNN_IP, rk.RK_KNN_L2):
        for sorted_ in (True, False):
            v, i, ws = rk.knn_topk(q, x, k, metric=metric, sorted=sorted_)
            assert rk.status(ws)[0] == 0
            _, uv, ui = _unfused(q, x, k, metric, sorted_)
            np.testing.assert_array_equal(i.cpu().numpy(), ui.cpu().numpy())
            np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32),
                                          uv.cpu().numpy().view(np.uint32))
