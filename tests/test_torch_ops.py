"""torch.ops.radik_amd.topk / topk_ragged: registration and shapes on the
CPU (fake tensors); on the GPU, identical to torch.topk on distinct values
(values and indices, sorted), the same set in index order when unsorted, and
the ragged form equal to the oracle with the tail fill."""

import numpy as np
import pytest


def test_ops_registered_with_schema():
    import torch

    import radik_amd.torch_ops  # noqa: F401
    s = str(torch.ops.radik_amd.topk.default._schema)
    assert "Tensor values, SymInt k, SymInt dim=-1, bool largest=True, bool sorted=True" in s
    assert "topk_ragged" in str(torch.ops.radik_amd.topk_ragged.default._schema)


def test_fake_shapes():
    import torch
    from torch._subclasses.fake_tensor import FakeTensorMode

    import radik_amd.torch_ops  # noqa: F401
    with FakeTensorMode():
        v, i = torch.ops.radik_amd.topk(torch.empty(3, 50, 7), 5, 1)
        assert v.shape == (3, 5, 7) and i.shape == (3, 5, 7) and i.dtype == torch.int64
        v, i = torch.ops.radik_amd.topk_ragged(torch.empty(100), torch.empty(5, dtype=torch.int64),
                                               torch.empty(4, dtype=torch.int32), 9)
        assert v.shape == (4, 9) and i.dtype == torch.int32


def _distinct(shape, dtype, seed):
    """Values distinct along the last dimension (a permutation of exactly
    representable integers, scaled)."""
    import torch
    g = torch.Generator().manual_seed(seed)
    n = shape[-1]
    rows = int(np.prod(shape[:-1])) if len(shape) > 1 else 1
    perm = torch.stack([torch.randperm(n, generator=g) for _ in range(rows)])
    x = (perm.to(torch.float32) - n / 2) / 4.0  # exact in fp32, f16 (n <= 2048), bf16 (n <= 256)
    return x.reshape(shape).to(dtype).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,n", [("float32", 5000), ("float16", 2048), ("bfloat16", 256)])
@pytest.mark.parametrize("largest", [True, False])
def test_topk_matches_torch_on_distinct_values(dtype, n, largest):
    import torch

    import radik_amd.torch_ops  # noqa: F401
    dt = getattr(torch, dtype)
    x = _distinct((37, n), dt, 1)
    for k in (1, 7, n // 3, n):
        rv, ri = torch.topk(x.float(), k, dim=-1, largest=largest, sorted=True)
        v, i = torch.ops.radik_amd.topk(x, k, -1, largest, True)
        assert i.dtype == torch.int64 and v.dtype == dt
        assert torch.equal(i, ri), (dtype, k, largest)
        assert torch.equal(v.float(), rv)
        # unsorted: the same elements, in index order
        uv, ui = torch.ops.radik_amd.topk(x, k, -1, largest, False)
        assert torch.equal(ui, torch.sort(ri, dim=-1).values)
        assert torch.equal(uv, torch.gather(x, -1, ui))


@pytest.mark.gpu
def test_topk_other_dims_and_shapes():
    import torch

    import radik_amd.torch_ops  # noqa: F401
    x = _distinct((4, 6, 1000), torch.float32, 2).transpose(1, 2)  # [4, 1000, 6], dim 1
    rv, ri = torch.topk(x, 50, dim=1)
    v, i = torch.ops.radik_amd.topk(x, 50, 1)
    assert v.shape == (4, 50, 6) and torch.equal(i, ri) and torch.equal(v, rv)
    y = _distinct((3000,), torch.float32, 3)
    rv, ri = torch.topk(y, 10, largest=False)
    v, i = torch.ops.radik_amd.topk(y, 10, 0, False)
    assert torch.equal(i, ri) and torch.equal(v, rv)
    v, i = torch.ops.radik_amd.topk(y, 0)
    assert v.numel() == 0 and i.numel() == 0
    with pytest.raises(ValueError):
        torch.ops.radik_amd.topk(y, 3001)


@pytest.mark.gpu
def test_topk_ties_lowest_index():
    import torch

    import radik_amd.torch_ops  # noqa: F401
    x = torch.zeros(2, 1000, device="cuda")
    x[:, 500:] = 1.0
    v, i = torch.ops.radik_amd.topk(x, 600, -1)
    expect = torch.cat([torch.arange(500, 1000), torch.arange(0, 100)]).cuda()
    assert torch.equal(i[0], expect) and torch.equal(i[1], expect)


@pytest.mark.gpu
def test_topk_ragged_matches_oracle():
    import torch

    import radik_amd.torch_ops  # noqa: F401
    import radik_amd.workloads as W
    from helpers import from_device_bits
    from oracle import oracle as orc
    w = W.c4("f32", scale=256, rows=300)
    vals = W.materialize(w)
    max_k = int(w.ks.max()) + 5
    v, i = torch.ops.radik_amd.topk_ragged(vals, torch.from_numpy(w.offsets()).cuda(),
                                           torch.from_numpy(w.ks).cuda(), max_k)
    oo = np.arange(w.rows + 1, dtype=np.int64) * max_k
    eb, ei = orc.topk(from_device_bits(vals, "f32")[: w.total], w.offsets(), w.ks,
                      flags=orc.LARGEST | orc.SORTED, out_off=oo, nthreads=8)
    np.testing.assert_array_equal(i.reshape(-1).cpu().numpy(), ei)
    np.testing.assert_array_equal(from_device_bits(v.reshape(-1), "f32"), eb)


@pytest.mark.gpu
def test_device_rejected_rows_read_empty_and_checked_raises():
    """A row the device rejects (its k above the max_k bound the host passed)
    is left at the pre-fill, value 0 / index -1, while the other rows are
    exact; checked=True reports the device status as RuntimeError."""
    import torch

    import radik_amd.torch_ops  # noqa: F401
    x = _distinct((3, 2000), torch.float32, 4).reshape(-1)
    off = torch.tensor([0, 2000, 4000, 6000], dtype=torch.int64, device="cuda")
    ks = torch.tensor([5, 50, 7], dtype=torch.int32, device="cuda")  # row 1: k > max_k
    v, i = torch.ops.radik_amd.topk_ragged(x, off, ks, 10)
    assert torch.equal(i[1], torch.full((10,), -1, dtype=torch.int32, device="cuda"))
    assert torch.equal(v[1], torch.zeros(10, device="cuda"))
    rv, ri = torch.topk(x[:2000], 5)
    assert torch.equal(i[0, :5].long(), ri) and torch.equal(i[0, 5:], i[1, 5:])
    rv, ri = torch.topk(x[4000:], 7)
    assert torch.equal(i[2, :7].long(), ri)
    with pytest.raises(RuntimeError, match="device status"):
        torch.ops.radik_amd.topk_ragged(x, off, ks, 10, checked=True)
    ks[1] = 9
    v2, i2 = torch.ops.radik_amd.topk_ragged(x, off, ks, 10, checked=True)
    assert torch.equal(i2[0], i[0]) and torch.equal(i2[2], i[2])
    v3, i3 = torch.ops.radik_amd.topk(x.reshape(3, 2000), 4, checked=True)
    assert torch.equal(i3, torch.topk(x.reshape(3, 2000), 4).indices)


@pytest.mark.gpu
def test_topk_refuses_a_dimension_above_int32():
    """Row positions are int32: a dimension of 2^31 elements is refused on
    the host (a stride-0 view, so nothing of that size is allocated)."""
    import torch

    import radik_amd.torch_ops  # noqa: F401
    big = torch.zeros(1, device="cuda").expand(2 ** 31)
    with pytest.raises(ValueError, match="2\\^31"):
        torch.ops.radik_amd.topk(big, 4)
