"""hipGraph capture of the batched call: a captured rk_topk_batched replays
correctly several times, with eager work and synchronizations between the
replays, and reads the buffers' current contents on each replay."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(scale=32, seed_shift=0):
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W
    w = W.c4("f32", scale=scale, rows=512)
    vals = W.materialize(w)
    d_off = torch.from_numpy(w.offsets()).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    oo = torch.from_numpy(w.out_offsets()).cuda()
    n = int(w.out_offsets()[-1])
    flags = rk.RK_LARGEST | rk.RK_SORTED
    ws = torch.empty(rk.workspace_size(rk.RK_F32, w.rows, w.total, w.max_k, flags),
                     dtype=torch.uint8, device="cuda")
    return w, vals, d_off, d_k, oo, n, ws


def test_graph_replays_match_eager():
    import torch

    import radik_amd as rk
    w, vals, d_off, d_k, oo, n, ws = _setup()
    ref_v = torch.empty(n, dtype=vals.dtype, device="cuda")
    ref_i = torch.empty(n, dtype=torch.int32, device="cuda")
    rk.topk_batched(vals, d_off, d_k, out_off=oo, out_vals=ref_v, out_idx=ref_i,
                    total_len=w.total, max_k_=w.max_k)
    torch.cuda.synchronize()

    out_v = torch.empty_like(ref_v)
    out_i = torch.empty_like(ref_i)
    s = torch.cuda.Stream()
    with rk.Graph.capture(s) as g:
        rk.topk_batched(vals, d_off, d_k, out_off=oo, out_vals=out_v, out_idx=out_i,
                        workspace=ws, total_len=w.total, max_k_=w.max_k, stream=s)
    for rep in range(4):
        out_v.fill_(0)
        out_i.fill_(-5)
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        assert torch.equal(out_i, ref_i), f"replay {rep}: indices differ"
        assert torch.equal(out_v.view(torch.int32), ref_v.view(torch.int32)), f"replay {rep}"
        # eager work between replays: the same call on the default stream with
        # its own workspace and outputs
        ev, ei, _, _ = rk.topk_batched(vals, d_off, d_k, total_len=w.total, max_k_=w.max_k)
        torch.cuda.synchronize()
        assert torch.equal(ei[:n], ref_i)
    g.close()


def test_graph_reads_current_inputs():
    import torch

    import radik_amd as rk
    from helpers import from_device_bits
    from oracle import oracle as orc
    w, vals, d_off, d_k, oo, n, ws = _setup(scale=64)
    out_v = torch.empty(n, dtype=vals.dtype, device="cuda")
    out_i = torch.empty(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with rk.Graph.capture(s) as g:
        rk.topk_batched(vals, d_off, d_k, out_off=oo, out_vals=out_v, out_idx=out_i,
                        workspace=ws, total_len=w.total, max_k_=w.max_k, stream=s)
    for seed in (11, 12, 13):
        rk.gen_fill(vals, rk.GEN_NORMAL, seed, 0)
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        bits = from_device_bits(vals, "f32")
        eb, ei = orc.topk(bits, w.offsets(), w.ks, flags=orc.LARGEST | orc.SORTED, nthreads=8)
        np.testing.assert_array_equal(out_i.cpu().numpy(), ei)
        np.testing.assert_array_equal(from_device_bits(out_v, "f32"), eb)
    g.close()
