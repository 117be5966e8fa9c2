"""PyTorch custom operators over the C ABI, so that PyTorch code can call
the MI355X top-k as ``torch.ops.radik_amd.*`` (registered on import):

  topk(values, k, dim=-1, largest=True, sorted=True) -> (values, int64 indices)
      torch.topk's signature and output: the k best of every slice along
      ``dim``; ``sorted`` orders them best first.  Ties at the k-th value are
      broken by the lower index and sorted ties by index ascending, so the
      result is deterministic; ``sorted=False`` returns index order.
  topk_ragged(values, offsets, ks, max_k, largest=True, sorted=True)
      -> (values [R, max_k], int32 indices [R, max_k])
      uneven rows of a 1-D tensor (row r = values[offsets[r]:offsets[r+1]],
      its own k = ks[r] <= max_k; offsets and ks on the device); row r fills
      min(len, k) slots and the rest with value 0 / index -1, as the
      reference's [taskNum][K] output (radik/RadixSelect/topk_radixselect.h:312-320).

Both run on the current stream without a host synchronisation (topk_ragged
takes the host bound max_k so that it needs none).  Everything the host can
check is checked before the launch (dtype, k against the dimension and the
sorted limit, a dimension or slice count above 2^31 - 1).  What only the
device sees (topk_ragged: a row whose k exceeds max_k, offsets that
decrease; an internal capacity failure) leaves that row unwritten: topk_ragged
pre-fills its outputs with value 0 / index -1, so such a row reads as empty.
``checked=True`` waits for the call and raises RuntimeError if the device
status of the call is not clean (one host synchronisation).
"""


import torch

from . import (RK_ASCEND, RK_LARGEST, RK_SORTED, _check, _dtype_code, lib, max_k as _max_k,
               status as _status, workspace_size)

_INT32_MAX = 2 ** 31 - 1

_SUPPORTED = (torch.float32, torch.float16, torch.bfloat16)


def _flags(largest: bool, sorted: bool) -> int:
    # best first, as torch.topk: descending for largest, ascending for smallest
    return (RK_LARGEST if largest else 0) | (RK_SORTED if sorted else 0) | (
        0 if largest else RK_ASCEND)


def _run(values, offsets, ks, num_rows, total, max_k, flags, out_v, out_i, out_off, checked,
         op):
    dt = _dtype_code(values)
    ws = torch.empty(workspace_size(dt, num_rows, total, max_k, flags), dtype=torch.uint8,
                     device=values.device)
    cur = torch.cuda.current_stream(values.device)
    # int64 index outputs are written directly (rk_topk_batched_i64)
    i64 = out_i.dtype == torch.int64
    fn = lib().rk_topk_batched_i64 if i64 else lib().rk_topk_batched
    rc = fn(dt, values.data_ptr(), offsets.data_ptr(), ks.data_ptr(), num_rows, total, max_k,
            flags, None, out_v.data_ptr(), out_i.data_ptr(), out_off.data_ptr(), ws.data_ptr(),
            ws.numel(), cur.cuda_stream)
    _check(rc, "rk_topk_batched_i64" if i64 else "rk_topk_batched")
    if checked:
        st = _status(ws, cur)
        if st[0] != 0:
            raise RuntimeError(f"radik_amd::{op}: device status 0x{st[0]:x} (bad row: 1, "
                               f"workspace bound: 2 or 4, internal: 8); failed rows are "
                               f"left unwritten")


@torch.library.custom_op("radik_amd::topk", mutates_args=(), device_types="cuda")
def topk(values: torch.Tensor, k: int, dim: int = -1, largest: bool = True,
         sorted: bool = True, checked: bool = False) -> tuple[torch.Tensor, torch.Tensor]:
    if values.dtype not in _SUPPORTED:
        raise TypeError(f"radik_amd::topk: unsupported dtype {values.dtype}")
    d = dim % max(values.dim(), 1)
    n = values.shape[d] if values.dim() else 1
    if n > _INT32_MAX:
        raise ValueError(f"radik_amd::topk: a dimension of size {n} exceeds 2^31 - 1 "
                         f"(row positions are int32)")
    if k < 0 or k > n:
        raise ValueError(f"radik_amd::topk: k={k} out of range for a dimension of size {n}")
    if sorted and k > _max_k():
        raise ValueError(f"radik_amd::topk: sorted output supports k up to {_max_k()} "
                         f"(k={k}); unsorted output has no limit")
    x = values.movedim(d, -1).contiguous() if values.dim() else values.reshape(1)
    lead = x.shape[:-1]
    rows = x.numel() // n if n else 0
    if rows >= 2 ** 31:
        raise ValueError("radik_amd::topk: more than 2^31 - 1 slices")
    out_v = torch.empty((*lead, k), dtype=values.dtype, device=values.device)
    out_i = torch.empty((*lead, k), dtype=torch.int64, device=values.device)
    if rows and k:
        dev = values.device
        off = torch.arange(0, rows + 1, dtype=torch.int64, device=dev) * n
        oo = torch.arange(0, rows + 1, dtype=torch.int64, device=dev) * k
        kk = torch.full((rows,), k, dtype=torch.int32, device=dev)
        _run(x, off, kk, rows, rows * n, k, _flags(largest, sorted), out_v, out_i, oo,
             checked, "topk")
    if values.dim():
        out_v, out_i = out_v.movedim(-1, d), out_i.movedim(-1, d)
    else:
        out_v, out_i = out_v.reshape(()), out_i.reshape(())
    return out_v.contiguous(), out_i.contiguous()


@topk.register_fake
def _topk_fake(values, k, dim=-1, largest=True, sorted=True, checked=False):
    shape = list(values.shape)
    if shape:
        shape[dim % len(shape)] = k
    return values.new_empty(shape), values.new_empty(shape, dtype=torch.int64)


@torch.library.custom_op("radik_amd::topk_ragged", mutates_args=(), device_types="cuda")
def topk_ragged(values: torch.Tensor, offsets: torch.Tensor, ks: torch.Tensor, max_k: int,
                largest: bool = True, sorted: bool = True,
                checked: bool = False) -> tuple[torch.Tensor, torch.Tensor]:
    if values.dtype not in _SUPPORTED or values.dim() != 1:
        raise TypeError("radik_amd::topk_ragged: values must be a 1-D float32/float16/bfloat16 "
                        "tensor")
    if offsets.dtype != torch.int64 or ks.dtype != torch.int32 or \
            offsets.numel() != ks.numel() + 1:
        raise TypeError("radik_amd::topk_ragged: offsets int64 [R+1] and ks int32 [R] expected")
    if max_k < 0 or (sorted and max_k > _max_k()):
        raise ValueError(f"radik_amd::topk_ragged: max_k={max_k} out of range for "
                         f"{'sorted' if sorted else 'unsorted'} output")
    rows = ks.numel()
    if rows > _INT32_MAX:
        raise ValueError("radik_amd::topk_ragged: more than 2^31 - 1 rows")
    # pre-filled (the tail fill), so a row the device rejects reads as empty
    out_v = torch.zeros((rows, max_k), dtype=values.dtype, device=values.device)
    out_i = torch.full((rows, max_k), -1, dtype=torch.int32, device=values.device)
    if rows and max_k:
        oo = torch.arange(0, rows + 1, dtype=torch.int64, device=values.device) * max_k
        _run(values.contiguous(), offsets.contiguous(), ks.contiguous(), rows, values.numel(),
             max_k, _flags(largest, sorted), out_v, out_i, oo, checked, "topk_ragged")
    return out_v, out_i


@topk_ragged.register_fake
def _topk_ragged_fake(values, offsets, ks, max_k, largest=True, sorted=True, checked=False):
    rows = ks.shape[0]
    return (values.new_empty((rows, max_k)),
            values.new_empty((rows, max_k), dtype=torch.int32))
