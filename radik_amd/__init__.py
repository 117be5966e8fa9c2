"""radik_amd — MI355X-native radix top-k (RadiK's batched top-k path).

Host-side mirror of the reference interface over the C-ABI library
``radik_amd/lib/libradik_amd.so`` (hand-written HIP for gfx950):

* :func:`topk_radix_select` mirrors ``topKRadixSelectL``
  (radik/RadixSelect/topk_radixselect.h:59-323): host ``taskLen``, one ``K``,
  output ``[taskNum][K]`` with the tail fill (index -1, value 0).
* :func:`get_workspace_size` mirrors ``getRadixSelectLWorkSpaceSize``
  (topk_radixselect.h:22-33).
* :func:`topk_batched` is the uneven-batch form: device row offsets, per-row
  k, per-row output offsets.

PyTorch is used only for device memory and streams.  There is no CPU
fallback: every call goes through the HIP library and raises if it is not
built or no GPU is present.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "RK_F32", "RK_F16", "RK_BF16", "RK_LARGEST", "RK_SORTED", "RK_ASCEND", "RK_EXACT",
    "RadikError", "lib", "lib_path", "max_k", "get_workspace_size", "workspace_size",
    "topk_radix_select", "topk_batched", "topk_rows", "status", "gen_fill",
    "get_tune", "set_tune", "timing_enable", "timing_read", "autotune_batched",
    "tune_cache_save", "tune_cache_load", "tune_cache_clear", "tune_cache_size", "Graph",
    "Server",
]

RK_F32, RK_F16, RK_BF16 = 0, 1, 2
RK_LARGEST, RK_SORTED, RK_ASCEND, RK_EXACT = 1, 2, 4, 8
RK_KNN_IP, RK_KNN_L2 = 0, 1
GEN_UNIFORM, GEN_NORMAL, GEN_ZIPF, GEN_NEARCONST, GEN_RANGE, GEN_CONST = range(6)
RK_OK, RK_E_ARG, RK_E_WORKSPACE, RK_E_UNSUPPORTED, RK_E_ALIGN, RK_E_HIP, RK_E_DEVICE = (
    0, -1, -2, -3, -4, -5, -6)
RK_E_IO, RK_E_SCRATCH = -7, -8
RK_NOALLOC_SCRATCH_BYTES = 64 << 20  # include/radik_amd.h

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("RADIK_AMD_LIB") or os.path.join(_HERE, "lib", "libradik_amd.so")
_lib = None

_ERRORS = {
    0: "ok", -1: "bad argument", -2: "workspace missing or too small",
    -3: "unsupported dtype, or sorted output with k above rk_max_k()", -4: "misaligned pointer",
    -5: "HIP runtime error", -6: "device-side validation failed",
    -7: "batch or result file missing, unreadable or malformed",
    -8: "call too large for the allocation-free entry's scratch",
}

# Every symbol include/radik_amd.h and include/radik_amd_harness.h declare.
EXPORTED = (
    "rk_max_k", "rk_topk_workspace_size", "rk_topk_batched", "rk_topk_batched_i64",
    "rk_topk_batched_noalloc", "rk_topk_rows", "rk_topk_compat_i64",
    "rk_compat_workspace_size", "rk_topk_compat", "rk_topk_status", "rk_strerror",
    "rk_get_tune", "rk_set_tune", "rk_gen_fill", "rk_timing_enable", "rk_timing_read",
    "rk_autotune_batched", "rk_tune_cache_save", "rk_tune_cache_load", "rk_tune_cache_clear",
    "rk_tune_cache_size", "rk_graph_begin", "rk_graph_end", "rk_graph_launch",
    "rk_graph_destroy", "rk_probe_read", "rk_server_create", "rk_server_submit",
    "rk_server_wait", "rk_server_drain", "rk_server_replay", "rk_server_get_stats",
    "rk_server_destroy", "rk_knn_workspace_size", "rk_knn_topk", "rk_knn_scores",
    "rk_batch_write", "rk_batch_info", "rk_batch_read", "rk_result_write", "rk_result_info",
    "rk_result_read",
)


class RadikError(RuntimeError):
    """A radik_amd C-ABI call returned an error code."""

    def __init__(self, code: int, where: str):
        super().__init__(f"{where}: {_ERRORS.get(code, 'unknown error')} (code {code})")
        self.code = code


class _Tune(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("tile", "grid", "sample_groups", "group_elems", "hist_bits", "unroll")]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch_keys", ctypes.c_int64), ("mem_budget", ctypes.c_int64),
                ("dtype", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("max_k", ctypes.c_int32), ("max_batch_rows", ctypes.c_int32),
                ("streams", ctypes.c_int32), ("plain", ctypes.c_int32)]


class _ServerStats(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int64) for n in
                 ("completed", "rejected", "failed", "batches", "fallback_batches",
                  "server_bytes", "peak_inflight_bytes")]
                + [(n, ctypes.c_double) for n in
                   ("mean_batch_rows", "p50_us", "p99_us", "mean_us", "max_us", "issue_us")]
                + [("max_batch_rows", ctypes.c_int32), ("pad_", ctypes.c_int32),
                   ("retained", ctypes.c_int64)])


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"radik_amd: {_LIB_PATH} is missing; build it with `python -m radik_amd.build`")
    # PyTorch-ROCm wheels carry their own HIP runtime, which torch loads by a
    # path of its own.  Loaded first, this library would bind the system
    # runtime, torch would then bring a second HIP/HSA runtime into the
    # process, and this library's would find no device.  With torch imported
    # first, the library binds torch's runtime (same soname): one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(_LIB_PATH)
    c = ctypes
    vp, i32, i64, u32, sz = c.c_void_p, c.c_int32, c.c_int64, c.c_uint32, c.c_size_t
    L.rk_max_k.restype = i32
    L.rk_topk_workspace_size.argtypes = [c.c_int, i32, i64, i32, u32, c.POINTER(sz)]
    L.rk_topk_batched.argtypes = [c.c_int, vp, vp, vp, i32, i64, i32, u32, vp, vp, vp, vp,
                                  vp, sz, vp]
    L.rk_topk_batched_i64.argtypes = [c.c_int, vp, vp, vp, i32, i64, i32, u32, vp, vp, vp, vp,
                                      vp, sz, vp]
    L.rk_topk_compat_i64.argtypes = [c.c_int, vp, vp, vp, vp, vp, sz, vp, i32, i32, u32, vp]
    L.rk_topk_batched_noalloc.argtypes = [c.c_int, vp, vp, vp, i32, i64, i32, u32, vp, vp, vp,
                                          vp, vp]
    L.rk_topk_rows.argtypes = [c.c_int, vp, vp, vp, i32, i64, i32, u32, vp, vp, vp, sz, vp]
    L.rk_compat_workspace_size.argtypes = [c.c_int, i32, i32, i32, u32, c.POINTER(sz)]
    L.rk_topk_compat.argtypes = [c.c_int, vp, vp, vp, vp, vp, sz, vp, i32, i32, u32, vp]
    L.rk_topk_status.argtypes = [vp, vp, c.POINTER(u32), c.POINTER(u32), c.POINTER(u32)]
    L.rk_strerror.restype = c.c_char_p
    L.rk_strerror.argtypes = [c.c_int]
    L.rk_get_tune.argtypes = [c.POINTER(_Tune)]
    L.rk_set_tune.argtypes = [c.POINTER(_Tune)]
    L.rk_gen_fill.argtypes = [c.c_int, c.c_int, c.c_uint64, c.c_uint64, i64, i64, c.c_double,
                              c.c_double, c.c_double, vp, vp]
    L.rk_autotune_batched.argtypes = [c.c_int, vp, vp, vp, i32, i64, i32, u32, vp, vp, vp, vp,
                                      sz, vp, i32, c.POINTER(_Tune), c.POINTER(c.c_float)]
    L.rk_tune_cache_save.argtypes = [c.c_char_p]
    L.rk_tune_cache_load.argtypes = [c.c_char_p]
    L.rk_graph_begin.argtypes = [vp]
    L.rk_graph_end.argtypes = [vp, c.POINTER(vp)]
    L.rk_graph_launch.argtypes = [vp, vp]
    L.rk_graph_destroy.argtypes = [vp]
    for name in ("rk_autotune_batched", "rk_tune_cache_save", "rk_tune_cache_load",
                 "rk_tune_cache_clear", "rk_tune_cache_size", "rk_graph_begin", "rk_graph_end",
                 "rk_graph_launch", "rk_graph_destroy"):
        getattr(L, name).restype = c.c_int
    L.rk_probe_read.argtypes = [vp, i64, i32, vp, vp]
    L.rk_probe_read.restype = c.c_int
    P = c.POINTER
    L.rk_server_create.argtypes = [P(_ServerConfig), P(vp)]
    L.rk_server_submit.argtypes = [vp, vp, i64, i32, vp, vp, P(c.c_uint64)]
    L.rk_server_wait.argtypes = [vp, c.c_uint64, P(c.c_double)]
    L.rk_server_drain.argtypes = [vp]
    L.rk_server_replay.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, P(c.c_double)]
    L.rk_server_get_stats.argtypes = [vp, P(_ServerStats)]
    L.rk_server_destroy.argtypes = [vp]
    for name in ("rk_server_create", "rk_server_submit", "rk_server_wait", "rk_server_drain",
                 "rk_server_replay", "rk_server_get_stats", "rk_server_destroy"):
        getattr(L, name).restype = c.c_int
    L.rk_knn_workspace_size.argtypes = [c.c_int, i32, i64, i32, i32, P(sz)]
    L.rk_knn_topk.argtypes = [c.c_int, c.c_int, vp, vp, i32, i64, i32, i32, u32, vp, vp, vp, sz,
                              vp]
    L.rk_knn_scores.argtypes = [c.c_int, c.c_int, vp, vp, i32, i64, i32, vp, vp]
    for name in ("rk_knn_workspace_size", "rk_knn_topk", "rk_knn_scores"):
        getattr(L, name).restype = c.c_int
    cp = c.c_char_p
    L.rk_batch_write.argtypes = [cp, c.c_int, u32, i32, vp, vp, vp]
    L.rk_batch_info.argtypes = [cp, vp]
    L.rk_batch_read.argtypes = [cp, i32, i64, i64, vp, vp, vp]
    L.rk_result_write.argtypes = [cp, c.c_int, i32, vp, vp, vp]
    L.rk_result_info.argtypes = [cp, vp]
    L.rk_result_read.argtypes = [cp, i32, i64, i64, vp, vp, vp]
    for name in ("rk_batch_write", "rk_batch_info", "rk_batch_read", "rk_result_write",
                 "rk_result_info", "rk_result_read"):
        getattr(L, name).restype = c.c_int
    L.rk_timing_enable.argtypes = [c.c_int]
    L.rk_timing_read.argtypes = [c.POINTER(c.c_double), c.POINTER(i32)]
    for name in ("rk_timing_enable", "rk_timing_read",
                 "rk_topk_workspace_size", "rk_topk_batched", "rk_topk_rows",
                 "rk_compat_workspace_size", "rk_topk_compat", "rk_topk_status",
                 "rk_get_tune", "rk_set_tune", "rk_gen_fill"):
        getattr(L, name).restype = c.c_int
    _lib = L
    return L


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise RadikError(rc, where)


def _torch():
    import torch
    return torch


def _dtype_code(t) -> int:
    torch = _torch()
    if t.dtype == torch.float32:
        return RK_F32
    if t.dtype == torch.float16:
        return RK_F16
    if t.dtype == torch.bfloat16:
        return RK_BF16
    raise TypeError(f"radik_amd: unsupported dtype {t.dtype}")


def _stream_ptr(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def _flags(largest: bool, sorted: bool, ascend: bool | None, exact: bool) -> int:
    if ascend is None:  # torch.topk order: best first
        ascend = not largest
    return ((RK_LARGEST if largest else 0) | (RK_SORTED if sorted else 0)
            | (RK_ASCEND if ascend else 0) | (RK_EXACT if exact else 0))


def max_k() -> int:
    """Largest k for sorted output (unsorted output accepts any k)."""
    return int(lib().rk_max_k())


def workspace_size(dtype: int, num_rows: int, total_len: int, max_k_: int, flags: int = 0) -> int:
    out = ctypes.c_size_t(0)
    _check(lib().rk_topk_workspace_size(dtype, num_rows, total_len, max_k_, flags,
                                        ctypes.byref(out)), "rk_topk_workspace_size")
    return int(out.value)


def get_workspace_size(K: int, n: int, taskNum: int, dtype: int = RK_F32) -> int:
    """Mirror of getRadixSelectLWorkSpaceSize<ValType>(K, n, taskNum, &bytes)."""
    out = ctypes.c_size_t(0)
    _check(lib().rk_compat_workspace_size(dtype, K, n, taskNum, 0, ctypes.byref(out)),
           "rk_compat_workspace_size")
    return int(out.value)


def _alloc_ws(nbytes: int, device):
    torch = _torch()
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)


def topk_radix_select(valIn, K: int, taskLen, *, idxIn=None, largest: bool = False,
                      ascend: bool = True, need_sorting: bool = True, exact: bool = False,
                      workspace=None, valOut=None, idxOut=None, stream=None,
                      idx_dtype=None):
    """Mirror of topKRadixSelectL<IdxType, LARGEST, ASCEND, ...>(valIn, idxIn,
    valOut, idxOut, workSpace, taskLen, taskNum, K, stream, needSorting).

    The reference's defaults in its harness are smallest-k, ascending
    (radik/test/test_radik.cu:18-20); they are kept here.  Returns
    ``(valOut, idxOut)`` shaped ``[taskNum, K]``.  WITHSCALE has no switch:
    the key-space scaling is always exact and always on.  IdxType is int32,
    or int64 (rk_topk_compat_i64) when ``idx_dtype`` is torch.int64 or an
    int64 ``idxIn`` / ``idxOut`` is given.
    """
    torch = _torch()
    taskLen = [int(x) for x in taskLen]
    taskNum = len(taskLen)
    dt = _dtype_code(valIn)
    if idx_dtype is None:
        idx_dtype = torch.int64 if any(t is not None and t.dtype == torch.int64
                                       for t in (idxIn, idxOut)) else torch.int32
    i64 = idx_dtype == torch.int64
    if valOut is None:
        valOut = torch.empty((taskNum, K), dtype=valIn.dtype, device=valIn.device)
    if idxOut is None:
        idxOut = torch.empty((taskNum, K), dtype=idx_dtype, device=valIn.device)
    if workspace is None:
        workspace = _alloc_ws(get_workspace_size(K, max(taskLen, default=0), taskNum, dt),
                              valIn.device)
    tl = (ctypes.c_int32 * max(taskNum, 1))(*taskLen)
    flags = _flags(largest, need_sorting, ascend, exact)
    fn = lib().rk_topk_compat_i64 if i64 else lib().rk_topk_compat
    rc = fn(dt, _ptr(valIn), _ptr(idxIn), _ptr(valOut), _ptr(idxOut), _ptr(workspace),
            workspace.numel(), tl, taskNum, K, flags, _stream_ptr(stream))
    _check(rc, "rk_topk_compat_i64" if i64 else "rk_topk_compat")
    return valOut, idxOut


def topk_batched(values, row_off, row_k, *, largest: bool = True, sorted: bool = True,
                 ascend: bool | None = None, exact: bool = False, idx_in=None, out_off=None,
                 total_len: int | None = None, max_k_: int | None = None, workspace=None,
                 out_vals=None, out_idx=None, stream=None, idx64: bool | None = None):
    """Uneven batched top-k.  ``values`` 1-D device tensor of concatenated rows,
    ``row_off`` int64 [R+1] and ``row_k`` int32 [R] on the device.  Output row
    r occupies ``[out_off[r], out_off[r+1])`` (default: packed by min(len, k)).
    Host bounds ``total_len`` / ``max_k_`` default to the exact values (which
    reads the metadata back; pass them to stay asynchronous).  Indices are
    int32, or int64 (rk_topk_batched_i64) when ``idx64`` is set or an int64
    ``idx_in`` / ``out_idx`` is given."""
    torch = _torch()
    if idx64 is None:
        idx64 = any(t is not None and t.dtype == torch.int64 for t in (idx_in, out_idx))
    dt = _dtype_code(values)
    R = row_k.numel()
    if total_len is None:
        total_len = int(row_off[-1].item() - row_off[0].item()) if R else 0
    if max_k_ is None:
        max_k_ = int(row_k.max().item()) if R else 0
    if out_off is None:
        lens = (row_off[1:] - row_off[:-1])
        m = torch.minimum(lens, row_k.to(torch.int64))
        out_off = torch.zeros(R + 1, dtype=torch.int64, device=values.device)
        out_off[1:] = torch.cumsum(m, 0)
    n_out = int(out_off[-1].item()) if out_vals is None or out_idx is None else 0
    if out_vals is None:
        out_vals = torch.empty(max(n_out, 1), dtype=values.dtype, device=values.device)
    idt = torch.int64 if idx64 else torch.int32
    if out_idx is None:
        out_idx = torch.empty(max(n_out, 1), dtype=idt, device=values.device)
    if out_idx.dtype != idt or (idx_in is not None and idx_in.dtype != idt):
        raise TypeError(f"topk_batched: index tensors must be {idt}")
    flags = _flags(largest, sorted, ascend, exact)
    if workspace is None:
        workspace = _alloc_ws(workspace_size(dt, R, total_len, max_k_, flags), values.device)
    fn = lib().rk_topk_batched_i64 if idx64 else lib().rk_topk_batched
    rc = fn(dt, _ptr(values), _ptr(row_off), _ptr(row_k), R, total_len, max_k_, flags,
            _ptr(idx_in), _ptr(out_vals), _ptr(out_idx), _ptr(out_off), _ptr(workspace),
            workspace.numel(), _stream_ptr(stream))
    _check(rc, "rk_topk_batched_i64" if idx64 else "rk_topk_batched")
    return out_vals, out_idx, out_off, workspace


def topk_rows(row_ptrs, row_len, row_k, out_val_ptrs, out_idx_ptrs, *, dtype: int,
              total_len: int, max_k_: int, largest: bool = True, sorted: bool = True,
              ascend: bool | None = None, exact: bool = False, workspace, stream=None) -> None:
    """Pointer form: each row and its k output slots at their own device
    addresses (device int64 tensors of pointers)."""
    flags = _flags(largest, sorted, ascend, exact)
    rc = lib().rk_topk_rows(dtype, _ptr(row_ptrs), _ptr(row_len), _ptr(row_k), row_k.numel(),
                            total_len, max_k_, flags, _ptr(out_val_ptrs), _ptr(out_idx_ptrs),
                            _ptr(workspace), workspace.numel(), _stream_ptr(stream))
    _check(rc, "rk_topk_rows")


def status(workspace, stream=None):
    """(status, fallback_rows, ordered_rows) of the last call on ``workspace``."""
    s, f, o = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    rc = lib().rk_topk_status(_ptr(workspace), _stream_ptr(stream), ctypes.byref(s),
                              ctypes.byref(f), ctypes.byref(o))
    if rc not in (0, -6):
        _check(rc, "rk_topk_status")
    return int(s.value), int(f.value), int(o.value)


def gen_fill(out, kind: int, seed: int, stream_id: int = 0, *, a: float = 0.0, b: float = 1.0,
             dup_frac: float = 0.0, row_len: int = 0, stream=None) -> None:
    """Seeded device-side input generation (harness; see radik_amd_harness.h)."""
    rc = lib().rk_gen_fill(_dtype_code(out), kind, seed, stream_id, out.numel(), row_len, a, b,
                           dup_frac, _ptr(out), _stream_ptr(stream))
    _check(rc, "rk_gen_fill")


def timing_enable(on: bool = True) -> None:
    _check(lib().rk_timing_enable(1 if on else 0), "rk_timing_enable")


def timing_read():
    """(summed filter-kernel milliseconds, launches timed) since the last read."""
    ms, n = ctypes.c_double(0.0), ctypes.c_int32(0)
    _check(lib().rk_timing_read(ctypes.byref(ms), ctypes.byref(n)), "rk_timing_read")
    return float(ms.value), int(n.value)


def autotune_batched(values, row_off, row_k, out_vals, out_idx, out_off, workspace, *,
                     total_len: int, max_k_: int, largest: bool = True, sorted: bool = True,
                     ascend: bool | None = None, exact: bool = False, reps: int = 3, stream=None):
    """Runs the autotuner on these buffers; returns (best tunables, median ms)."""
    t = _Tune()
    ms = ctypes.c_float(0.0)
    flags = _flags(largest, sorted, ascend, exact)
    rc = lib().rk_autotune_batched(_dtype_code(values), _ptr(values), _ptr(row_off), _ptr(row_k),
                                   row_k.numel(), total_len, max_k_, flags, _ptr(out_vals),
                                   _ptr(out_idx), _ptr(out_off), _ptr(workspace),
                                   workspace.numel(), _stream_ptr(stream), reps, ctypes.byref(t),
                                   ctypes.byref(ms))
    _check(rc, "rk_autotune_batched")
    return {n: getattr(t, n) for n, _ in _Tune._fields_}, float(ms.value)


def tune_cache_save(path: str) -> None:
    _check(lib().rk_tune_cache_save(path.encode()), "rk_tune_cache_save")


def tune_cache_load(path: str) -> None:
    _check(lib().rk_tune_cache_load(path.encode()), "rk_tune_cache_load")


def tune_cache_clear() -> None:
    _check(lib().rk_tune_cache_clear(), "rk_tune_cache_clear")


def tune_cache_size() -> int:
    """Number of shape buckets with an autotuned entry."""
    return int(lib().rk_tune_cache_size())


class Graph:
    """A hipGraph of radik_amd calls captured on a (non-null) stream:

        with rk.Graph.capture(stream) as g:
            rk.topk_batched(..., stream=stream)
        g.launch(stream)
    """

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def capture(cls, stream):
        return _GraphCapture(stream)

    def launch(self, stream=None) -> None:
        _check(lib().rk_graph_launch(self._h, _stream_ptr(stream)), "rk_graph_launch")

    def close(self) -> None:
        if self._h:
            lib().rk_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _GraphCapture:
    def __init__(self, stream):
        self.stream = stream
        self.graph = None

    def __enter__(self):
        _check(lib().rk_graph_begin(_stream_ptr(self.stream)), "rk_graph_begin")
        self.graph = Graph(None)
        return self.graph

    def __exit__(self, et, ev, tb):
        h = ctypes.c_void_p()
        rc = lib().rk_graph_end(_stream_ptr(self.stream), ctypes.byref(h))
        if et is None:
            _check(rc, "rk_graph_end")
            self.graph._h = h.value
        return False


def knn_workspace_size(dtype: int, nq: int, n: int, d: int, k: int) -> int:
    out = ctypes.c_size_t(0)
    _check(lib().rk_knn_workspace_size(dtype, nq, n, d, k, ctypes.byref(out)),
           "rk_knn_workspace_size")
    return out.value


def knn_topk(queries, base, k: int, *, metric: int = RK_KNN_IP, sorted: bool = True,
             workspace=None, out_vals=None, out_idx=None, stream=None):
    """Fused k-nearest-neighbour select: the k best database rows of every
    query (``queries`` [nq, d], ``base`` [n, d], bf16 or f16 device tensors)
    by inner product (largest) or squared L2 distance (smallest), without the
    [nq, n] score matrix in memory.  Returns (values [nq, k] float32, indices
    [nq, k] int32, workspace)."""
    torch = _torch()
    if queries.dim() != 2 or base.dim() != 2 or queries.shape[1] != base.shape[1]:
        raise ValueError("knn_topk: queries [nq, d] and base [n, d] expected")
    if queries.dtype != base.dtype:
        raise ValueError("knn_topk: queries and base must share a dtype")
    dt = _dtype_code(queries)
    nq, d = queries.shape
    n = base.shape[0]
    if workspace is None:
        workspace = _alloc_ws(knn_workspace_size(dt, nq, n, d, k), queries.device)
    if out_vals is None:
        out_vals = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
    if out_idx is None:
        out_idx = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
    rc = lib().rk_knn_topk(dt, metric, _ptr(queries.contiguous()), _ptr(base.contiguous()), nq, n,
                           d, k, RK_SORTED if sorted else 0, _ptr(out_vals), _ptr(out_idx),
                           _ptr(workspace), workspace.numel(), _stream_ptr(stream))
    _check(rc, "rk_knn_topk")
    return out_vals, out_idx, workspace


def knn_scores(queries, base, *, metric: int = RK_KNN_IP, out=None, stream=None):
    """The [nq, n] float32 scores rk_knn_topk selects from, written to memory
    (same arithmetic: the unfused path)."""
    torch = _torch()
    dt = _dtype_code(queries)
    nq, d = queries.shape
    n = base.shape[0]
    if out is None:
        out = torch.empty((nq, n), dtype=torch.float32, device=queries.device)
    rc = lib().rk_knn_scores(dt, metric, _ptr(queries.contiguous()), _ptr(base.contiguous()), nq,
                             n, d, _ptr(out), _stream_ptr(stream))
    _check(rc, "rk_knn_scores")
    return out


class _FileInfo(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("flags", ctypes.c_uint32), ("rows", ctypes.c_int64),
                ("total", ctypes.c_int64)]


_HOST_BITS = {RK_F32: np.uint32, RK_F16: np.uint16, RK_BF16: np.uint16}


def _host(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return a, a.ctypes.data_as(ctypes.c_void_p)


def save_batch(path: str, dtype: int, values_bits, row_off, row_k, flags: int = 0) -> None:
    """Writes a CSR batch file (rk_batch_write).  ``values_bits``: host array of
    the values' bit patterns (uint32 for RK_F32, uint16 for RK_F16 / RK_BF16);
    ``row_off`` int64 [R+1] from 0; ``row_k`` int32 [R]; ``flags`` the RK_*
    flags the batch is meant for."""
    v, vp = _host(values_bits, _HOST_BITS[dtype])
    o, op = _host(row_off, np.int64)
    k, kp = _host(row_k, np.int32)
    _check(lib().rk_batch_write(path.encode(), dtype, flags, len(k), op, kp, vp), "rk_batch_write")


def load_batch(path: str):
    """Reads a CSR batch file: (dtype, flags, values bits, row_off, row_k) as
    host numpy arrays."""
    info = _FileInfo()
    _check(lib().rk_batch_info(path.encode(), ctypes.byref(info)), "rk_batch_info")
    v = np.empty(info.total, dtype=_HOST_BITS[info.dtype])
    o = np.empty(info.rows + 1, dtype=np.int64)
    k = np.empty(info.rows, dtype=np.int32)
    _check(lib().rk_batch_read(path.encode(), info.dtype, info.rows, info.total,
                               o.ctypes.data_as(ctypes.c_void_p),
                               k.ctypes.data_as(ctypes.c_void_p),
                               v.ctypes.data_as(ctypes.c_void_p)), "rk_batch_read")
    return info.dtype, info.flags, v, o, k


def save_results(path: str, dtype: int, out_off, out_vals_bits, out_idx) -> None:
    """Writes a result file (rk_result_write) from host arrays."""
    o, op = _host(out_off, np.int64)
    v, vp = _host(out_vals_bits, _HOST_BITS[dtype])
    i, ip = _host(out_idx, np.int32)
    _check(lib().rk_result_write(path.encode(), dtype, len(o) - 1, op, vp, ip), "rk_result_write")


def load_results(path: str):
    """Reads a result file: (dtype, out_off, out_vals bits, out_idx)."""
    info = _FileInfo()
    _check(lib().rk_result_info(path.encode(), ctypes.byref(info)), "rk_result_info")
    o = np.empty(info.rows + 1, dtype=np.int64)
    v = np.empty(info.total, dtype=_HOST_BITS[info.dtype])
    i = np.empty(info.total, dtype=np.int32)
    _check(lib().rk_result_read(path.encode(), info.dtype, info.rows, info.total,
                                o.ctypes.data_as(ctypes.c_void_p),
                                v.ctypes.data_as(ctypes.c_void_p),
                                i.ctypes.data_as(ctypes.c_void_p)), "rk_result_read")
    return info.dtype, o, v, i


def probe_read(t, blocks_per_cu: int = 0, stream=None) -> None:
    """Launches the streaming-read probe over tensor ``t`` (harness) with
    ``blocks_per_cu`` workgroups per CU (0: the default)."""
    torch = _torch()
    out = torch.zeros(1, dtype=torch.int32, device=t.device)
    _check(lib().rk_probe_read(_ptr(t), t.numel() * t.element_size(), blocks_per_cu, _ptr(out),
                               _stream_ptr(stream)), "rk_probe_read")


def get_tune() -> dict:
    t = _Tune()
    _check(lib().rk_get_tune(ctypes.byref(t)), "rk_get_tune")
    return {n: getattr(t, n) for n, _ in _Tune._fields_}


def set_tune(**kw) -> None:
    """Sets the named tunables (0 = that field's default); others keep their value."""
    t = _Tune(**{n: int(kw.get(n, -1)) for n, _ in _Tune._fields_})
    _check(lib().rk_set_tune(ctypes.byref(t)), "rk_set_tune")


class Server:
    """The continuous-batching serving loop (rk_server_*).

    Requests are device rows (tensors, or raw device pointers with lengths)
    and their own output buffers; ``replay`` drives an open-loop arrival
    schedule from the C++ side and returns per-request latencies (us).
    """

    def __init__(self, dtype=RK_F32, *, largest=True, sorted=True, ascend=False, max_k=4096,
                 max_batch_rows=1024, max_batch_keys=1 << 26, streams=4,
                 mem_budget=32 << 30, plain=False):
        flags = ((RK_LARGEST if largest else 0) | (RK_SORTED if sorted else 0)
                 | (RK_ASCEND if ascend else 0))
        cfg = _ServerConfig(max_batch_keys=int(max_batch_keys), mem_budget=int(mem_budget),
                            dtype=int(dtype), flags=flags, max_k=int(max_k),
                            max_batch_rows=int(max_batch_rows), streams=int(streams),
                            plain=1 if plain else 0)
        h = ctypes.c_void_p()
        _check(lib().rk_server_create(ctypes.byref(cfg), ctypes.byref(h)), "rk_server_create")
        self._h = h.value

    def submit(self, vals, k: int, out_vals, out_idx, length: int | None = None) -> int:
        """Queues top-k of the device row ``vals``; returns the ticket."""
        n = vals.numel() if length is None else int(length)
        t = ctypes.c_uint64()
        _check(lib().rk_server_submit(self._h, _ptr(vals), n, int(k), _ptr(out_vals),
                                      _ptr(out_idx), ctypes.byref(t)), "rk_server_submit")
        return t.value

    def wait(self, ticket: int) -> float:
        lat = ctypes.c_double()
        _check(lib().rk_server_wait(self._h, int(ticket), ctypes.byref(lat)), "rk_server_wait")
        return lat.value

    def wait_code(self, ticket: int) -> int:
        """Return code of the request (0, or an RK_E_* code) instead of raising."""
        return lib().rk_server_wait(self._h, int(ticket), None)

    def drain(self) -> None:
        _check(lib().rk_server_drain(self._h), "rk_server_drain")

    def replay(self, val_ptrs, lens, ks, out_val_ptrs, out_idx_ptrs, arrivals):
        """Submits request i at arrivals[i] s (open loop); returns (latency_us, makespan_s)."""
        n = len(lens)
        P = np.ascontiguousarray(np.asarray(val_ptrs, dtype=np.uint64))
        L = np.ascontiguousarray(np.asarray(lens, dtype=np.int64))
        K = np.ascontiguousarray(np.asarray(ks, dtype=np.int32))
        OV = np.ascontiguousarray(np.asarray(out_val_ptrs, dtype=np.uint64))
        OI = np.ascontiguousarray(np.asarray(out_idx_ptrs, dtype=np.uint64))
        A = np.ascontiguousarray(np.asarray(arrivals, dtype=np.float64))
        lat = np.zeros(n, np.float64)
        mk = ctypes.c_double()
        _check(lib().rk_server_replay(self._h, n, P.ctypes.data, L.ctypes.data, K.ctypes.data,
                                      OV.ctypes.data, OI.ctypes.data, A.ctypes.data,
                                      lat.ctypes.data, ctypes.byref(mk)), "rk_server_replay")
        return lat, mk.value

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib().rk_server_get_stats(self._h, ctypes.byref(st)), "rk_server_get_stats")
        return {n: getattr(st, n) for n, _ in _ServerStats._fields_ if n != "pad_"}

    def close(self) -> None:
        if self._h:
            lib().rk_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
