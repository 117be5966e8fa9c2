"""ORACLE loader — test infrastructure only.

Thin ctypes wrapper over ``oracle/_build/liboracle.so`` (built from
``oracle/oracle.cpp`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may use
it, and only as the checker / the reported CPU baseline.  The product
package ``radik_amd`` never imports this module.

The semantics (restated from radik/RadixSelect/topk_radixselect.h:59-323 and
radik/test/test_radik.cu:188-213) are documented in oracle.cpp's header.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

DT_F32, DT_F16, DT_BF16 = 0, 1, 2
LARGEST, SORTED, ASCEND = 1, 2, 4


def build() -> str:
    """Compile the oracle with its Makefile (g++ only, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_flip.restype = ctypes.c_uint32
        L.orc_flip.argtypes = [ctypes.c_int, ctypes.c_uint32]
        L.orc_sel_key.restype = ctypes.c_uint32
        L.orc_sel_key.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int]
        L.orc_topk.restype = ctypes.c_int
        L.orc_topk.argtypes = [
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
        ]
        L.orc_hardware_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def flip(dt: int, bits: int) -> int:
    return int(lib().orc_flip(dt, bits))


def sel_key(dt: int, bits: int, largest: bool) -> int:
    return int(lib().orc_sel_key(dt, bits, 1 if largest else 0))


def hardware_threads() -> int:
    return int(lib().orc_hardware_threads())


def usable_threads() -> int:
    """Threads the multithreaded oracle may use here: the CPUs in this
    process's affinity mask, bounded by the cgroup CPU quota when one is set
    (a container's share of a larger host), else all hardware threads."""
    n = hardware_threads()
    try:
        n = min(n, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        pass
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            with open(path) as f:
                quota, period = f.read().split()[:2]
            if quota != "max":
                n = min(n, max(1, -(-int(quota) // int(period))))
        except (OSError, ValueError):
            pass
    return max(1, n)


def _ptr(a):
    return None if a is None else a.ctypes.data


def topk(bits: np.ndarray, row_off, row_k=None, K: int = -1, *, dt: int = DT_F32,
         flags: int = LARGEST | SORTED, idx_in=None, out_off=None,
         nthreads: int = 1):
    """Batched top-k over raw bit patterns.

    ``bits``: uint32 (f32) or uint16 (f16/bf16) array of concatenated rows.
    Returns ``(out_bits, out_idx)`` laid out as ``out_off`` (or ``[R][K]``).
    """
    bits = np.ascontiguousarray(bits)
    want = np.uint32 if dt == DT_F32 else np.uint16
    if bits.dtype != want:
        raise TypeError(f"bits must be {want}")
    row_off = np.ascontiguousarray(row_off, dtype=np.int64)
    R = len(row_off) - 1
    rk = None if row_k is None else np.ascontiguousarray(row_k, dtype=np.int32)
    if out_off is None and K < 0:
        # default layout: rows packed by min(len, k)
        lens = row_off[1:] - row_off[:-1]
        out_off = np.zeros(R + 1, np.int64)
        out_off[1:] = np.cumsum(np.minimum(lens, rk.astype(np.int64)))
    oo = None if out_off is None else np.ascontiguousarray(out_off, dtype=np.int64)
    total_out = int(oo[-1]) if oo is not None else R * K
    out_bits = np.zeros(max(total_out, 1), dtype=want)
    out_idx = np.zeros(max(total_out, 1), dtype=np.int32)
    ii = None if idx_in is None else np.ascontiguousarray(idx_in, dtype=np.int32)
    rc = lib().orc_topk(dt, _ptr(bits), _ptr(row_off), _ptr(rk), K, R, flags,
                        _ptr(ii), _ptr(out_bits), _ptr(out_idx), _ptr(oo),
                        nthreads)
    if rc != 0:
        raise ValueError("orc_topk: bad argument")
    return out_bits[:total_out], out_idx[:total_out]
