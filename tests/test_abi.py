"""CPU tests of the C-ABI library: it loads, exports every symbol the public
headers declare, and its host-side logic (workspace sizing, argument
validation, tunables) behaves; no kernel is launched here."""

from __future__ import annotations

import ctypes
import glob
import os
import re
import subprocess

import pytest

import radik_amd as rk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(rk.lib_path()):
        from radik_amd import build
        build.build(verbose=False)
    return rk.lib()


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(rk_\w+)\s*\(", text))
    return names


def test_headers_declare_the_expected_entry_points():
    assert _declared_symbols() == set(rk.EXPORTED)


def test_library_exports_every_declared_symbol(L):
    out = subprocess.run(["nm", "-D", "--defined-only", rk.lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = _declared_symbols() - exported
    assert not missing, f"not exported: {missing}"
    for name in _declared_symbols():
        assert hasattr(L, name)


def test_library_is_built_for_gfx950(L):
    # the offload bundle of the library names its target
    with open(rk.lib_path(), "rb") as f:
        assert b"gfx950" in f.read()


def test_workspace_size_grows_with_bounds(L):
    base = rk.workspace_size(rk.RK_F32, 16, 1 << 20, 256)
    assert base > 0
    assert rk.workspace_size(rk.RK_F32, 32, 1 << 20, 256) > base
    assert rk.workspace_size(rk.RK_F32, 16, 1 << 24, 256) > base
    assert rk.workspace_size(rk.RK_F32, 16, 1 << 20, 4096) > base
    # uneven batch of 2048 rows over 3.5 G keys: bounded by total length,
    # not rows x longest row (which would be 2048 * 2^24 * 8 B = 275 GB)
    big = rk.workspace_size(rk.RK_F32, 2048, 3_540_000_000, 4096)
    assert big < 2 << 30
    assert rk.workspace_size(rk.RK_BF16, 2048, 3_540_000_000, 4096) == big


def test_compat_workspace_size_matches_reference_arguments(L):
    # getRadixSelectLWorkSpaceSize(K, n, taskNum): n is the longest row
    a = rk.get_workspace_size(1024, 1 << 17, 512)
    b = rk.workspace_size(rk.RK_F32, 512, 512 << 17, 1024)
    assert a >= b


def test_size_query_rejects_bad_arguments(L):
    sz = ctypes.c_size_t()
    assert L.rk_topk_workspace_size(9, 1, 1, 1, 0, ctypes.byref(sz)) == -3
    assert L.rk_topk_workspace_size(0, -1, 1, 1, 0, ctypes.byref(sz)) == -1
    assert L.rk_topk_workspace_size(0, 1, -5, 1, 0, ctypes.byref(sz)) == -1
    assert L.rk_topk_workspace_size(0, 1, 1, 1, 0, None) == -1
    assert L.rk_compat_workspace_size(0, -1, 1, 1, 0, ctypes.byref(sz)) == -1


def test_batched_rejects_bad_arguments_before_touching_the_device(L):
    # unknown dtype, negative sizes, missing metadata / outputs
    args = [None] * 3
    assert L.rk_topk_batched(7, *args, 1, 10, 1, 0, None, None, None, None, None, 0, None) == -3
    assert L.rk_topk_batched(0, *args, -1, 10, 1, 0, None, None, None, None, None, 0, None) == -1
    assert L.rk_topk_batched(0, 1, 2, None, 1, 10, 1, 0, None, 3, 4, 5, None, 0, None) == -1
    assert L.rk_topk_batched(0, 1, 2, 3, 1, 10, 1, 0, None, None, 4, 5, None, 0, None) == -1
    # zero rows is a no-op
    assert L.rk_topk_batched(0, 1, 2, 3, 0, 0, 0, 0, None, 4, 5, 6, None, 0, None) == 0
    # compat: null workspace is an error (topk_radixselect.h:77-79)
    tl = (ctypes.c_int32 * 1)(10)
    assert L.rk_topk_compat(0, 1, None, 2, 3, None, 0, tl, 1, 4, 0, None) == -2
    assert L.rk_topk_compat(0, 1, None, 2, 3, 8, 0, None, 1, 4, 0, None) == -1


def test_knn_host_side_validation(L):
    sz = ctypes.c_size_t()
    assert L.rk_knn_workspace_size(rk.RK_BF16, 1024, 1_000_000, 128, 100, ctypes.byref(sz)) == 0
    big = sz.value
    assert L.rk_knn_workspace_size(rk.RK_BF16, 2048, 1_000_000, 128, 100, ctypes.byref(sz)) == 0
    assert sz.value > big  # grows with the queries
    assert L.rk_knn_workspace_size(rk.RK_F32, 4, 100, 64, 5, ctypes.byref(sz)) == -3
    assert L.rk_knn_workspace_size(rk.RK_F16, 4, 100, 40, 5, ctypes.byref(sz)) == -1  # d % 16
    assert L.rk_knn_workspace_size(rk.RK_F16, 4, 100, 272, 5, ctypes.byref(sz)) == -1  # d > 256
    assert L.rk_knn_workspace_size(rk.RK_F16, 4, 100, 64, 101, ctypes.byref(sz)) == -1  # k > n
    assert L.rk_knn_workspace_size(rk.RK_F16, 4, 100, 64, 0, ctypes.byref(sz)) == -1
    # rk_knn_topk / rk_knn_scores validate before any launch
    assert L.rk_knn_topk(rk.RK_BF16, 2, 16, 32, 4, 100, 64, 5, 2, 48, 64, 256, 1 << 20,
                         None) == -1  # unknown metric
    assert L.rk_knn_topk(rk.RK_BF16, 0, 16, 32, 4, 100, 64, 5, 2, 48, 64, None, 0, None) == -2
    assert L.rk_knn_topk(rk.RK_BF16, 0, 17, 32, 4, 100, 64, 5, 2, 48, 64, 256, 1 << 20,
                         None) == -4  # queries not 16-byte aligned
    assert L.rk_knn_topk(rk.RK_BF16, 0, 16, 32, 4, 100, 64, 5, 2, 48, 64, 256, 16,
                         None) == -2  # workspace too small
    assert L.rk_knn_topk(rk.RK_BF16, 0, 16, 32, 4, 100, 64, 5, 1, 48, 64, 256, 1 << 20,
                         None) == -1  # flags other than RK_SORTED
    assert L.rk_knn_scores(rk.RK_BF16, 0, 16, 32, 4, 100, 64, None, None) == -1


def test_strerror(L):
    for code in (0, -1, -2, -3, -4, -5, -6, 42):
        assert L.rk_strerror(code)


def test_tune_roundtrip(L):
    old = rk.get_tune()
    try:
        rk.set_tune(tile=4096, grid=512, unroll=8)
        t = rk.get_tune()
        assert (t["tile"], t["grid"], t["unroll"]) == (4096, 512, 8)
        assert t["hist_bits"] == old["hist_bits"]
    finally:
        rk.set_tune(**old)


def test_product_never_imports_the_oracle():
    """The product package must not reach the checker."""
    for path in glob.glob(os.path.join(ROOT, "radik_amd", "**", "*.*"), recursive=True):
        if path.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
            text = open(path).read()
            assert "oracle" not in text.replace("the oracle", ""), path


def test_server_argument_validation():
    """The serving entry points reject bad configurations and null handles
    before touching the device."""
    import ctypes

    import radik_amd as rk
    L = rk.lib()
    h = ctypes.c_void_p()
    cfg = rk._ServerConfig(max_batch_keys=1 << 20, mem_budget=1 << 30, dtype=rk.RK_F32,
                           flags=rk.RK_LARGEST, max_k=64, max_batch_rows=0, streams=1, plain=0)
    assert L.rk_server_create(ctypes.byref(cfg), ctypes.byref(h)) == rk.RK_E_ARG
    cfg.max_batch_rows = 8
    cfg.flags = rk.RK_LARGEST | rk.RK_EXACT  # the exact path is chosen by `plain`
    assert L.rk_server_create(ctypes.byref(cfg), ctypes.byref(h)) == rk.RK_E_ARG
    cfg.flags = rk.RK_LARGEST
    cfg.dtype = 7
    assert L.rk_server_create(ctypes.byref(cfg), ctypes.byref(h)) == rk.RK_E_UNSUPPORTED
    cfg.dtype = rk.RK_F32
    cfg.mem_budget = 0
    assert L.rk_server_create(ctypes.byref(cfg), ctypes.byref(h)) == rk.RK_E_ARG
    assert L.rk_server_create(None, ctypes.byref(h)) == rk.RK_E_ARG
    t = ctypes.c_uint64()
    assert L.rk_server_submit(None, None, 0, 0, None, None, ctypes.byref(t)) == rk.RK_E_ARG
    assert L.rk_server_wait(None, 0, None) == rk.RK_E_ARG
    assert L.rk_server_drain(None) == rk.RK_E_ARG
    assert L.rk_server_get_stats(None, None) == rk.RK_E_ARG
    assert L.rk_server_destroy(None) == rk.RK_E_ARG
    assert ctypes.sizeof(rk._ServerConfig) == 40 and ctypes.sizeof(rk._ServerStats) == 120
