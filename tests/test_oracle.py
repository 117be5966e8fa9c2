"""CPU tests pinning the oracle (oracle/oracle.cpp) before it is trusted as
the checker: known answers, exhaustive 16-bit checks, the independent
brute-force golden vectors, closed-form special cases and invariants."""

from __future__ import annotations

import struct

import numpy as np
import pytest

from oracle import oracle as orc
from helpers import DT_CODE, case_arrays, load_golden


def test_flip_known_answers():
    doc = load_golden()
    for bits, key in doc["flip_kats_f32"]:
        assert orc.flip(orc.DT_F32, bits) == key, hex(bits)
    # smallest-k uses the complement order
    assert orc.sel_key(orc.DT_F32, 0x3F800000, False) == (~0xBF800000) & 0xFFFFFFFF
    # NaN is the worst key in both directions
    for nan in (0x7FC00000, 0xFFC00000, 0x7F800001):
        assert orc.sel_key(orc.DT_F32, nan, True) == 0
        assert orc.sel_key(orc.DT_F32, nan, False) == 0


def _decode16(dt: str, bits: np.ndarray) -> np.ndarray:
    if dt == "f16":
        return bits.astype(np.uint16).view(np.float16).astype(np.float64)
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("dt", ["f16", "bf16"])
def test_flip_16bit_exhaustive(dt):
    """All 2^16 patterns: key order == numeric order (-0 < +0), NaN excluded;
    and the 16-bit key order equals the order of the exact fp32 widening."""
    code = DT_CODE[dt]
    bits = np.arange(1 << 16, dtype=np.uint32)
    keys = np.array([orc.flip(code, int(b)) for b in bits], dtype=np.uint32)
    assert len(np.unique(keys)) == 1 << 16  # a bijection
    vals = _decode16(dt, bits)
    ok = ~np.isnan(vals)
    order = np.argsort(keys[ok], kind="stable")
    v = vals[ok][order]
    assert np.all(np.diff(v) >= 0)
    # -0 sorts directly below +0
    b_sorted = bits[ok][order]
    zero_pos = np.nonzero(v == 0)[0]
    assert len(zero_pos) == 2
    assert b_sorted[zero_pos[0]] == 0x8000 and b_sorted[zero_pos[1]] == 0x0000
    # widening: the fp32 keys of the widened values are in the same order
    f32_bits = vals[ok].astype(np.float32).view(np.uint32)
    k32 = np.array([orc.flip(orc.DT_F32, int(b)) for b in f32_bits], dtype=np.uint64)
    assert np.all(np.diff(k32[order].astype(np.int64)) > 0)


def test_flip_f32_random_vs_numeric():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, size=20000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32).astype(np.float64)
    ok = ~np.isnan(vals)
    bits, vals = bits[ok], vals[ok]
    keys = np.array([orc.flip(orc.DT_F32, int(b)) for b in bits], dtype=np.uint64)
    order = np.argsort(keys)
    assert np.all(np.diff(vals[order]) >= 0)


@pytest.mark.parametrize("ci", range(72))
def test_oracle_matches_golden(ci):
    doc = load_golden()
    case = doc["cases"][ci]
    bits, off, ks, oo, exp_bits, exp_idx = case_arrays(case)
    got_bits, got_idx = orc.topk(bits, off, ks, dt=DT_CODE[case["dtype"]], flags=case["flags"],
                                 out_off=oo)
    np.testing.assert_array_equal(got_bits, exp_bits)
    np.testing.assert_array_equal(got_idx, exp_idx)


def _f32_bits(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x.astype(np.float32)).view(np.uint32)


def test_k1_is_argmax_argmin():
    rng = np.random.default_rng(1)
    x = rng.integers(-50, 50, size=(16, 300)).astype(np.float32)  # many ties, no -0
    x[x == 0] = 1
    bits = _f32_bits(x.ravel())
    off = np.arange(0, 16 * 300 + 1, 300, dtype=np.int64)
    for largest in (True, False):
        ob, oi = orc.topk(bits, off, np.ones(16, np.int32), flags=orc.SORTED | (orc.LARGEST if largest else 0))
        ref = x.argmax(1) if largest else x.argmin(1)  # first occurrence = lowest index
        np.testing.assert_array_equal(oi, ref)
        np.testing.assert_array_equal(ob.view(np.float32), x[np.arange(16), ref])


def test_k_eq_len_is_stable_sort():
    rng = np.random.default_rng(2)
    x = rng.integers(1, 20, size=(4, 257)).astype(np.float32)
    bits = _f32_bits(x.ravel())
    off = np.arange(0, 4 * 257 + 1, 257, dtype=np.int64)
    ks = np.full(4, 257, np.int32)
    _, oi = orc.topk(bits, off, ks, flags=orc.SORTED | orc.LARGEST)
    ref = np.argsort(-x, axis=1, kind="stable")
    np.testing.assert_array_equal(oi.reshape(4, 257), ref)
    _, oi = orc.topk(bits, off, ks, flags=orc.SORTED | orc.ASCEND)
    ref = np.argsort(x, axis=1, kind="stable")
    np.testing.assert_array_equal(oi.reshape(4, 257), ref)


def test_all_equal_takes_lowest_positions():
    bits = _f32_bits(np.full(1000, 2.5))
    off = np.array([0, 1000], np.int64)
    for flags in (orc.LARGEST | orc.SORTED, orc.SORTED, orc.LARGEST, orc.SORTED | orc.ASCEND):
        _, oi = orc.topk(bits, off, np.array([37], np.int32), flags=flags)
        np.testing.assert_array_equal(oi, np.arange(37))


def test_matches_torch_topk_on_distinct_values():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    R, n = 8, 5000
    x = rng.permutation(R * n).astype(np.float32).reshape(R, n) * 0.25 - 1000.0
    bits = _f32_bits(x.ravel())
    off = np.arange(0, R * n + 1, n, dtype=np.int64)
    for largest in (True, False):
        ob, oi = orc.topk(bits, off, np.full(R, 64, np.int32),
                          flags=orc.SORTED | (orc.LARGEST if largest else orc.ASCEND))
        tv, ti = torch.topk(torch.from_numpy(x), 64, dim=1, largest=largest, sorted=True)
        np.testing.assert_array_equal(ob.view(np.float32).reshape(R, 64), tv.numpy())
        np.testing.assert_array_equal(oi.reshape(R, 64), ti.numpy())


def test_invariants_with_duplicates_and_tail_fill():
    rng = np.random.default_rng(4)
    lens = rng.integers(0, 3000, size=30)
    ks = rng.integers(0, 400, size=30).astype(np.int32)
    off = np.zeros(31, np.int64)
    off[1:] = np.cumsum(lens)
    x = rng.integers(-30, 30, size=int(off[-1])).astype(np.float32) * 0.5 + 0.25
    bits = _f32_bits(x)
    K = 400
    ob, oi = orc.topk(bits, off, ks, K=K, flags=orc.LARGEST)  # unsorted, [R][K] layout
    ob = ob.view(np.float32).reshape(30, K)
    oi = oi.reshape(30, K)
    for r in range(30):
        row = x[off[r]:off[r + 1]]
        m = min(len(row), ks[r])
        sel = oi[r, :m]
        assert np.all(np.diff(sel) > 0)  # unsorted output is in position order
        np.testing.assert_array_equal(ob[r, :m], row[sel])
        assert np.all(oi[r, m:] == -1) and np.all(ob[r, m:] == 0)
        if m and m < len(row):
            kth = row[sel].min()
            rest = np.setdiff1d(np.arange(len(row)), sel)
            assert np.all(row[rest] <= kth)
            # ties at the k-th value: the selected ones have the lowest positions
            eq_all = np.nonzero(row == kth)[0]
            eq_sel = sel[row[sel] == kth]
            np.testing.assert_array_equal(eq_sel, eq_all[: len(eq_sel)])


def test_multithreaded_equals_single_threaded():
    rng = np.random.default_rng(5)
    off = np.arange(0, 64 * 4096 + 1, 4096, dtype=np.int64)
    bits = _f32_bits(rng.standard_normal(64 * 4096))
    ks = rng.integers(1, 300, size=64).astype(np.int32)
    a = orc.topk(bits, off, ks, flags=orc.LARGEST | orc.SORTED, nthreads=1)
    b = orc.topk(bits, off, ks, flags=orc.LARGEST | orc.SORTED, nthreads=8)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_bf16_f16_widening_consistency():
    """A 16-bit row gives the same positions as its exact fp32 widening."""
    rng = np.random.default_rng(6)
    for dt in ("f16", "bf16"):
        b16 = rng.integers(0, 1 << 16, size=5000).astype(np.uint16)
        vals = _decode16(dt, b16)
        b16 = b16[~np.isnan(vals)]
        vals = vals[~np.isnan(vals)]
        off = np.array([0, len(b16)], np.int64)
        b32 = vals.astype(np.float32).view(np.uint32)
        for flags in (orc.LARGEST | orc.SORTED, orc.SORTED):
            _, i16 = orc.topk(b16, off, np.array([300], np.int32), dt=DT_CODE[dt], flags=flags)
            _, i32 = orc.topk(b32, off, np.array([300], np.int32), flags=flags)
            np.testing.assert_array_equal(i16, i32)


def test_struct_decode_sanity():
    # guards the fixture generator's decoding conventions
    assert struct.unpack("<e", struct.pack("<H", 0x3C00))[0] == 1.0
    assert struct.unpack("<f", struct.pack("<I", 0x3F80 << 16))[0] == 1.0
