"""Shared helpers of the test suite: fixture loading and running the HIP
path (through the C-ABI) on raw bit patterns, so results compare bit-exactly
with the oracle and the golden vectors."""

from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "topk_golden.json")

DT_CODE = {"f32": 0, "f16": 1, "bf16": 2}
NP_BITS = {"f32": np.uint32, "f16": np.uint16, "bf16": np.uint16}


def load_golden():
    with open(GOLDEN) as f:
        return json.load(f)


def case_arrays(case):
    """Concatenated bits, row offsets, ks and output offsets of a golden case."""
    dt = case["dtype"]
    rows = case["rows"]
    lens = [len(r) for r in rows]
    off = np.zeros(len(rows) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    bits = np.array([b for r in rows for b in r], dtype=NP_BITS[dt])
    ks = np.array(case["ks"], np.int32)
    oo = np.zeros(len(rows) + 1, np.int64)
    oo[1:] = np.cumsum(case["slots"])
    exp_bits = np.array([b for r in case["expect_bits"] for b in r], dtype=NP_BITS[dt])
    exp_idx = np.array([i for r in case["expect_idx"] for i in r], dtype=np.int32)
    return bits, off, ks, oo, exp_bits, exp_idx


def _torch_dtype(dt):
    import torch
    return {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dt]


def to_device_bits(bits: np.ndarray, dt: str, pad_front: int = 0):
    """Device tensor holding `bits`, optionally preceded by pad_front
    elements (to misalign row starts relative to 16-byte boundaries)."""
    import torch
    raw = np.ascontiguousarray(bits)
    if pad_front:
        raw = np.concatenate([np.zeros(pad_front, raw.dtype), raw])
    sview = raw.view(np.int32 if dt == "f32" else np.int16)
    t = torch.from_numpy(sview.copy()).cuda()
    return t.view(_torch_dtype(dt))[pad_front:]


def from_device_bits(t, dt: str) -> np.ndarray:
    import torch
    iv = t.view(torch.int32 if dt == "f32" else torch.int16).cpu().numpy()
    return iv.view(NP_BITS[dt])


def gpu_topk(dt: str, bits: np.ndarray, off: np.ndarray, ks: np.ndarray, flags: int,
             out_off: np.ndarray | None = None, idx_in: np.ndarray | None = None,
             pad_front: int = 0, check_status: bool = True):
    """Runs rk_topk_batched on the GPU; returns (out_bits, out_idx, status tuple)."""
    import torch

    import radik_amd as rk

    if out_off is None:
        lens = off[1:] - off[:-1]
        m = np.minimum(lens, ks.astype(np.int64))
        out_off = np.zeros(len(ks) + 1, np.int64)
        out_off[1:] = np.cumsum(m)
    R = len(ks)
    vals = to_device_bits(bits, dt, pad_front)
    d_off = torch.from_numpy(off.astype(np.int64)).cuda()
    d_k = torch.from_numpy(ks.astype(np.int32)).cuda()
    d_oo = torch.from_numpy(out_off.astype(np.int64)).cuda()
    n_out = int(out_off[-1])
    ov = torch.full((max(n_out, 1),), 0, dtype=_torch_dtype(dt), device="cuda")
    oi = torch.full((max(n_out, 1),), -7, dtype=torch.int32, device="cuda")
    ii = None if idx_in is None else torch.from_numpy(idx_in.astype(np.int32)).cuda()
    total = int(off[-1] - off[0])
    mk = int(ks.max()) if R else 0
    ws_bytes = rk.workspace_size(DT_CODE[dt], R, total, mk, flags)
    ws = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device="cuda")
    L = rk.lib()
    rc = L.rk_topk_batched(DT_CODE[dt], vals.data_ptr(), d_off.data_ptr(), d_k.data_ptr(), R,
                           total, mk, flags, None if ii is None else ii.data_ptr(),
                           ov.data_ptr(), oi.data_ptr(), d_oo.data_ptr(), ws.data_ptr(),
                           ws.numel(), torch.cuda.current_stream().cuda_stream)
    if rc != 0:
        raise rk.RadikError(rc, "rk_topk_batched")
    st = rk.status(ws)
    if check_status:
        assert st[0] == 0, f"device status {st}"
    torch.cuda.synchronize()
    return from_device_bits(ov, dt)[:n_out], oi.cpu().numpy()[:n_out], st
