"""CSR batch and result files (rk_batch_* / rk_result_*): host-only, so
tested on the CPU: round trips for every dtype (ragged and empty rows, an
empty batch), the header fields, and rejection of malformed files and
arguments."""

from __future__ import annotations

import os

import numpy as np
import pytest

import radik_amd as rk

DTYPES = [(rk.RK_F32, np.uint32), (rk.RK_F16, np.uint16), (rk.RK_BF16, np.uint16)]


def _batch(seed, dt_bits, rows=7):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 50, rows)
    lens[2] = 0  # an empty row
    off = np.zeros(rows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    ks = rng.integers(0, 20, rows).astype(np.int32)
    bits = rng.integers(0, np.iinfo(dt_bits).max, int(off[-1]), dtype=np.uint64).astype(dt_bits)
    return bits, off, ks


@pytest.mark.parametrize("dtype,bits_t", DTYPES)
def test_batch_round_trip(tmp_path, dtype, bits_t):
    bits, off, ks = _batch(dtype + 3, bits_t)
    path = str(tmp_path / "b.rkb")
    flags = rk.RK_LARGEST | rk.RK_SORTED
    rk.save_batch(path, dtype, bits, off, ks, flags=flags)
    dt, fl, v, o, k = rk.load_batch(path)
    assert dt == dtype and fl == flags
    assert v.dtype == bits_t
    np.testing.assert_array_equal(v, bits)
    np.testing.assert_array_equal(o, off)
    np.testing.assert_array_equal(k, ks)
    # header 64 + offsets + k padded to 8 + values
    kb = 4 * len(ks)
    assert os.path.getsize(path) == 64 + 8 * len(off) + kb + (8 - kb % 8) % 8 + bits.nbytes


def test_empty_batch(tmp_path):
    path = str(tmp_path / "e.rkb")
    rk.save_batch(path, rk.RK_F32, np.zeros(0, np.uint32), np.zeros(1, np.int64),
                  np.zeros(0, np.int32))
    dt, fl, v, o, k = rk.load_batch(path)
    assert dt == rk.RK_F32 and len(v) == 0 and list(o) == [0] and len(k) == 0


@pytest.mark.parametrize("dtype,bits_t", DTYPES)
def test_result_round_trip(tmp_path, dtype, bits_t):
    rng = np.random.default_rng(dtype)
    m = rng.integers(0, 9, 5)
    off = np.zeros(6, np.int64)
    off[1:] = np.cumsum(m)
    vals = rng.integers(0, 1 << 16, int(off[-1])).astype(bits_t)
    idx = rng.integers(-1, 1000, int(off[-1])).astype(np.int32)
    path = str(tmp_path / "r.rkr")
    rk.save_results(path, dtype, off, vals, idx)
    dt, o, v, i = rk.load_results(path)
    assert dt == dtype
    np.testing.assert_array_equal(o, off)
    np.testing.assert_array_equal(v, vals)
    np.testing.assert_array_equal(i, idx)


def _code(fn):
    with pytest.raises(rk.RadikError) as e:
        fn()
    return e.value.code


def test_malformed_files_rejected(tmp_path):
    bits, off, ks = _batch(1, np.uint32)
    good = str(tmp_path / "g.rkb")
    rk.save_batch(good, rk.RK_F32, bits, off, ks)
    raw = open(good, "rb").read()
    cases = {
        "missing": None,
        "magic": b"XKBATCH1" + raw[8:],
        "truncated": raw[:-3],
        "trailing": raw + b"\0",
        "dtype": raw[:8] + (9).to_bytes(4, "little") + raw[12:],
    }
    # offsets that decrease: swap row_off[1] and row_off[3] (both nonzero)
    o = np.frombuffer(raw[64:64 + 8 * len(off)], np.int64).copy()
    o[1], o[3] = o[3], o[1]
    cases["offsets"] = raw[:64] + o.tobytes() + raw[64 + 8 * len(off):]
    for name, data in cases.items():
        p = str(tmp_path / f"{name}.rkb")
        if data is not None:
            open(p, "wb").write(data)
        assert _code(lambda: rk.load_batch(p)) == -7, name
    # a result file is not a batch file, and vice versa
    assert _code(lambda: rk.load_results(good)) == -7


def test_bad_arguments(tmp_path):
    p = str(tmp_path / "x.rkb")
    bits = np.zeros(4, np.uint32)
    ks = np.ones(2, np.int32)
    # offsets not starting at 0 / decreasing
    assert _code(lambda: rk.save_batch(p, rk.RK_F32, bits, [1, 2, 4], ks)) == -1
    assert _code(lambda: rk.save_batch(p, rk.RK_F32, bits, [0, 3, 2], ks)) == -1
    # negative k
    assert _code(lambda: rk.save_batch(p, rk.RK_F32, bits, [0, 2, 4], [1, -1])) == -1
    # unknown dtype through the C ABI
    L = rk.lib()
    off = np.array([0, 4], np.int64)
    k1 = np.ones(1, np.int32)
    rc = L.rk_batch_write(p.encode(), 7, 0, 1, off.ctypes.data_as(rk.ctypes.c_void_p),
                          k1.ctypes.data_as(rk.ctypes.c_void_p),
                          bits.ctypes.data_as(rk.ctypes.c_void_p))
    assert rc == -3
    assert L.rk_batch_info(None, None) == -1


def test_read_refuses_a_file_larger_than_the_buffers(tmp_path):
    """*_read takes the sizes the caller allocated for (from *_info): a file
    replaced by a larger one in between is refused before anything is
    written, for both file kinds."""
    L = rk.lib()
    vp = rk.ctypes.c_void_p
    small, big = str(tmp_path / "s.rkb"), str(tmp_path / "b.rkb")
    rk.save_batch(small, rk.RK_F32, np.arange(4, dtype=np.uint32), [0, 4], [2])
    rk.save_batch(big, rk.RK_F32, np.arange(40, dtype=np.uint32), [0, 10, 40], [2, 3])
    off = np.full(2, -5, np.int64)  # sized for the small file: 1 row, 4 values
    ks = np.full(1, -5, np.int32)
    v = np.full(4, 7, np.uint32)
    args = (off.ctypes.data_as(vp), ks.ctypes.data_as(vp), v.ctypes.data_as(vp))
    assert L.rk_batch_read(big.encode(), rk.RK_F32, 1, 4, *args) == -7
    assert list(off) == [-5, -5] and list(ks) == [-5] and list(v) == [7] * 4
    assert L.rk_batch_read(small.encode(), rk.RK_F32, 1, 4, *args) == 0
    assert list(off) == [0, 4] and list(ks) == [2] and list(v) == [0, 1, 2, 3]
    assert L.rk_batch_read(small.encode(), rk.RK_F32, -1, 4, *args) == -1

    rs, rb = str(tmp_path / "s.rkr"), str(tmp_path / "b.rkr")
    rk.save_results(rs, rk.RK_F32, [0, 2], np.ones(2, np.uint32), np.zeros(2, np.int32))
    rk.save_results(rb, rk.RK_F32, [0, 2, 9], np.ones(9, np.uint32), np.zeros(9, np.int32))
    ro = np.full(2, -5, np.int64)
    rv = np.full(2, 7, np.uint32)
    ri = np.full(2, 7, np.int32)
    rargs = (ro.ctypes.data_as(vp), rv.ctypes.data_as(vp), ri.ctypes.data_as(vp))
    assert L.rk_result_read(rb.encode(), rk.RK_F32, 1, 2, *rargs) == -7
    assert list(ro) == [-5, -5] and list(rv) == [7, 7]
    assert L.rk_result_read(rs.encode(), rk.RK_F32, 1, 2, *rargs) == 0
    assert list(ro) == [0, 2] and list(rv) == [1, 1] and list(ri) == [0, 0]


def test_read_refuses_a_file_whose_dtype_changed(tmp_path):
    """A file rewritten with 4-byte values after *_info reported 2-byte ones
    (same rows and totals) would overflow buffers sized for the old dtype:
    *_read takes the dtype it sized for and refuses the file untouched."""
    L = rk.lib()
    vp = rk.ctypes.c_void_p
    p = str(tmp_path / "x.rkb")
    rk.save_batch(p, rk.RK_BF16, np.arange(4, dtype=np.uint16), [0, 4], [2])
    info = rk._FileInfo()
    assert L.rk_batch_info(p.encode(), rk.ctypes.byref(info)) == 0
    assert info.dtype == rk.RK_BF16
    rk.save_batch(p, rk.RK_F32, np.arange(4, dtype=np.uint32), [0, 4], [2])  # swapped
    off = np.full(2, -5, np.int64)
    ks = np.full(1, -5, np.int32)
    v = np.full(4 + 8, 7, np.uint16)  # sized for 4 bf16 values, plus a guard
    args = (off.ctypes.data_as(vp), ks.ctypes.data_as(vp), v.ctypes.data_as(vp))
    assert L.rk_batch_read(p.encode(), info.dtype, info.rows, info.total, *args) == -7
    assert list(off) == [-5, -5] and list(ks) == [-5] and list(v) == [7] * 12
    assert L.rk_batch_read(p.encode(), rk.RK_F32, 1, 4, *args) == 0
    assert L.rk_batch_read(p.encode(), 9, 1, 4, *args) == -3

    r = str(tmp_path / "x.rkr")
    rk.save_results(r, rk.RK_F16, [0, 2], np.ones(2, np.uint16), np.zeros(2, np.int32))
    assert L.rk_result_info(r.encode(), rk.ctypes.byref(info)) == 0
    rk.save_results(r, rk.RK_F32, [0, 2], np.ones(2, np.uint32), np.zeros(2, np.int32))
    ro = np.full(2, -5, np.int64)
    rv = np.full(2 + 4, 7, np.uint16)
    ri = np.full(2, 7, np.int32)
    rargs = (ro.ctypes.data_as(vp), rv.ctypes.data_as(vp), ri.ctypes.data_as(vp))
    assert L.rk_result_read(r.encode(), info.dtype, info.rows, info.total, *rargs) == -7
    assert list(ro) == [-5, -5] and list(rv) == [7] * 6 and list(ri) == [7, 7]
