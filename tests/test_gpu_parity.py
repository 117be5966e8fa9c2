"""GPU parity: the HIP path (through the C-ABI) against the golden vectors and
the oracle, bit-exact on values and indices, on seeded inputs covering every
path of the pipeline (copy-all, LDS-direct, sampled filter, exact passes,
ordered emit), every dtype and flag combination, ragged / misaligned rows,
ties, NaN / inf / signed zeros, and the reference-surface compat entry."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import DT_CODE, NP_BITS, case_arrays, from_device_bits, gpu_topk, load_golden
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LARGEST, SORTED, ASCEND, EXACT = 1, 2, 4, 8
FLAG_SETS = [LARGEST | SORTED, SORTED | ASCEND, LARGEST | SORTED | ASCEND, SORTED, LARGEST, 0]


def _oracle(dt, bits, off, ks, flags, idx_in=None):
    return orc.topk(bits, off, ks, dt=DT_CODE[dt], flags=flags & 7, idx_in=idx_in, nthreads=8)


def _check(dt, bits, off, ks, flags, pad_front=0, idx_in=None, expect_clean=True):
    gb, gi, st = gpu_topk(dt, bits, off, ks, flags, idx_in=idx_in, pad_front=pad_front)
    ob, oi = _oracle(dt, bits, off, ks, flags, idx_in=idx_in)
    if not np.array_equal(gi, oi) or not np.array_equal(gb, ob):
        bad = np.nonzero((gi != oi) | (gb != ob))[0]
        raise AssertionError(f"{dt} flags={flags} mismatch at {bad[:10]} of {len(gi)}; "
                             f"gpu idx {gi[bad[:5]]} oracle idx {oi[bad[:5]]} status {st}")
    return st


def _gen_bits(dt, kind, n, seed, a=0.0, b=1.0, dup=0.0, row_len=0):
    import torch

    import radik_amd as rk
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dt]
    t = torch.empty(max(n, 1), dtype=tdt, device="cuda")
    rk.gen_fill(t[:n], kind, seed, 0, a=a, b=b, dup_frac=dup, row_len=row_len)
    torch.cuda.synchronize()
    return from_device_bits(t, dt)[:n]


@pytest.mark.parametrize("ci", range(72))
@pytest.mark.parametrize("exact", [False, True])
def test_golden_vectors(ci, exact):
    case = load_golden()["cases"][ci]
    bits, off, ks, oo, exp_bits, exp_idx = case_arrays(case)
    dt = case["dtype"]
    gb, gi, _ = gpu_topk(dt, bits, off, ks, case["flags"] | (EXACT if exact else 0), out_off=oo)
    np.testing.assert_array_equal(gi, exp_idx)
    np.testing.assert_array_equal(gb, exp_bits)


def _ragged(seed, nrows, lo, hi):
    rng = np.random.default_rng(seed)
    lens = np.floor(2.0 ** rng.uniform(np.log2(lo), np.log2(hi), nrows)).astype(np.int64)
    off = np.zeros(nrows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    return rng, lens, off


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("flags", FLAG_SETS)
@pytest.mark.parametrize("exact", [False, True])
def test_ragged_normal_rows(dt, flags, exact):
    rng, lens, off = _ragged(11 + DT_CODE[dt], 24, 300, 400_000)
    ks = np.minimum(np.floor(2.0 ** rng.uniform(0, 12, len(lens))), 4096).astype(np.int32)
    bits = _gen_bits(dt, 1, int(off[-1]), 100 + flags)
    _check(dt, bits, off, ks, flags | (EXACT if exact else 0))


@pytest.mark.parametrize("pad", [1, 2, 3])
def test_misaligned_rows(pad):
    rng, lens, off = _ragged(21, 12, 17000, 90_000)
    lens[::3] += 1  # odd lengths shift every later row start
    off[1:] = np.cumsum(lens)
    ks = rng.integers(1, 2000, len(lens)).astype(np.int32)
    bits = _gen_bits("f32", 1, int(off[-1]), 7)
    for exact in (False, True):
        _check("f32", bits, off, ks, LARGEST | SORTED | (EXACT if exact else 0), pad_front=pad)
    b16 = _gen_bits("bf16", 1, int(off[-1]), 8)
    _check("bf16", b16, off, ks, SORTED | ASCEND, pad_front=pad)


@pytest.mark.parametrize("dt", ["f32", "bf16"])
@pytest.mark.parametrize("pad", [0, 1, 5])
def test_row_heads_and_tails_at_every_alignment(dt, pad):
    # The filter streams 16-byte vectors from each row's first element on a
    # 128-byte boundary; the elements before it (up to 63) and after the
    # last whole vector are read one per lane.  64 rows of 30001 elements
    # start at every residue mod 64 elements, and each row's best keys are
    # planted in its first 64 and last 8 elements (ties, broken by index).
    rows, n = 64, 30_001
    off = np.arange(rows + 1, dtype=np.int64) * n
    ks = np.full(rows, 72, np.int32)
    bits = _gen_bits(dt, 1, rows * n, 31 + pad)
    to_bits = (lambda v: np.float32(v).view(np.uint32)) if dt == "f32" else \
        (lambda v: np.uint16(np.float32(v).view(np.uint32) >> 16))
    for flags, head, tail in ((LARGEST | SORTED, 1000.0, 2000.0),
                              (SORTED | ASCEND, -1000.0, -2000.0)):
        b = bits.copy()
        for r in range(rows):
            b[off[r]:off[r] + 64] = to_bits(head)
            b[off[r + 1] - 8:off[r + 1]] = to_bits(tail)
        _check(dt, b, off, ks, flags, pad_front=pad)


@pytest.mark.parametrize("exact", [False, True])
def test_all_equal_long_rows_take_lowest_positions(exact):
    """Every key ties: the ties overflow the candidate slots and the ordered
    emit must return positions 0..k-1 (a row of 2^18 equal values)."""
    n = 1 << 18
    off = np.array([0, n, 2 * n, 3 * n], np.int64)
    bits = np.concatenate([np.full(n, 0x40200000, np.uint32), np.zeros(n, np.uint32),
                           np.full(n, 0xBF800000, np.uint32)])
    ks = np.array([1000, 16384, 1], np.int32)
    for flags in (LARGEST | SORTED, SORTED | ASCEND, LARGEST):
        st = _check("f32", bits, off, ks, flags | (EXACT if exact else 0))
        assert st[2] >= 2  # rows 0 and 1 went through the ordered emit


@pytest.mark.parametrize("largest", [True, False])
def test_near_constant_with_duplicates(largest):
    n = 1 << 19
    off = np.arange(0, 4 * n + 1, n, dtype=np.int64)
    bits = _gen_bits("f32", 3, 4 * n, 31, dup=0.3, row_len=n)
    ks = np.array([16384, 4096, 100, 1], np.int32)
    for flags in ((LARGEST if largest else ASCEND) | SORTED, LARGEST if largest else 0):
        for exact in (False, True):
            _check("f32", bits, off, ks, flags | (EXACT if exact else 0))


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
def test_exact_passes_skip_shared_key_bits(dt):
    """The exact passes skip the bits all counted keys share once one bin
    holds them all: rows whose keys agree on their top 12..31 bits (a few
    low bits vary, with ties), a constant row (every bit shared: the k-th
    key is resolved by the skip alone), a row whose values differ only in
    the last bit, and a negative near-constant row, in both directions and
    every order, on the exact path and (for comparison) the sampled path."""
    rng = np.random.default_rng(77 + DT_CODE[dt])
    dtype = np.uint32 if dt == "f32" else np.uint16
    rows = []
    base = {"f32": 0x3F80_0000, "bf16": 0x3F80, "f16": 0x3C00}[dt]  # 1.0
    nbase = {"f32": 0xBF80_0000, "bf16": 0xBF80, "f16": 0xBC00}[dt]  # -1.0
    for vary in ((2, 5, 9, 13) if dt == "f32" else (2, 4, 7)):
        rows.append((base + rng.integers(0, 1 << vary, 40_000)).astype(dtype))
    rows.append(np.full(30_000, base, dtype))                         # constant
    rows.append((base + rng.integers(0, 2, 50_000)).astype(dtype))    # last bit only
    rows.append((nbase + rng.integers(0, 1 << 6, 45_000)).astype(dtype))
    lens = np.array([len(r) for r in rows], np.int64)
    off = np.zeros(len(rows) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    bits = np.concatenate(rows)
    ks = np.array([100, 1000, 3000, 17, 500, 2000, 8000][:len(rows)], np.int32)
    for flags in (LARGEST | SORTED, SORTED | ASCEND, LARGEST, 0, LARGEST | SORTED | ASCEND):
        for exact in (EXACT, 0):
            _check(dt, bits, off, ks, flags | exact)


@pytest.mark.parametrize("largest", [True, False])
def test_zipf_rows(largest):
    n = 1 << 19
    off = np.arange(0, 3 * n + 1, n, dtype=np.int64)
    bits = _gen_bits("f32", 2, 3 * n, 41, a=1.1, b=2.0 ** 20)
    ks = np.array([16384, 3000, 7], np.int32)
    flags = (LARGEST if largest else ASCEND) | SORTED
    for exact in (False, True):
        _check("f32", bits, off, ks, flags | (EXACT if exact else 0))


@pytest.mark.parametrize("dt", ["f32", "bf16"])
def test_special_values(dt):
    rng = np.random.default_rng(5)
    n = 70_000
    bits = _gen_bits(dt, 1, 4 * n, 9, a=0.0, b=2.0)
    specials = {"f32": [0x7FC00000, 0xFFC00000, 0x7F800000, 0xFF800000, 0, 0x80000000],
                "bf16": [0x7FC0, 0xFFC0, 0x7F80, 0xFF80, 0, 0x8000]}[dt]
    pos = rng.integers(0, 4 * n, 20000)
    bits[pos] = np.array(specials, dtype=NP_BITS[dt])[rng.integers(0, 6, 20000)]
    off = np.arange(0, 4 * n + 1, n, dtype=np.int64)
    ks = np.array([4096, 1, 300, 16384], np.int32)
    for flags in FLAG_SETS:
        for exact in (False, True):
            _check(dt, bits, off, ks, flags | (EXACT if exact else 0))


def test_nan_heavy_row_selects_nan_last():
    n = 40_000
    bits = np.full(n, 0x7FC00000, np.uint32)
    bits[::10] = np.arange(0, n, 10, dtype=np.float32).view(np.uint32)
    off = np.array([0, n], np.int64)
    for exact in (False, True):
        _check("f32", bits, off, np.array([6000], np.int32), LARGEST | SORTED | (EXACT if exact else 0))


def test_idx_in_gathered():
    rng, lens, off = _ragged(51, 8, 100, 200_000)
    ks = rng.integers(1, 3000, len(lens)).astype(np.int32)
    bits = _gen_bits("f32", 1, int(off[-1]), 52)
    idx_in = rng.integers(-2**31, 2**31 - 1, int(off[-1])).astype(np.int32)
    for flags in (LARGEST | SORTED, SORTED | ASCEND, 0):
        _check("f32", bits, off, ks, flags, idx_in=idx_in)


def test_k_edges_and_empty_rows():
    lens = np.array([0, 1, 5, 16384, 16385, 100_000, 1 << 20, 3], np.int64)
    ks = np.array([5, 1, 0, 16384, 16384, 1, 16384, 3], np.int32)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    bits = _gen_bits("f32", 1, int(off[-1]), 61)
    for flags in (LARGEST | SORTED, SORTED | ASCEND):
        for exact in (False, True):
            _check("f32", bits, off, ks, flags | (EXACT if exact else 0))


def test_compat_surface_tail_fill():
    """topk_radix_select mirrors topKRadixSelectL: [taskNum][K] output, one K,
    rows shorter than K tail-filled with index -1 / value 0."""
    import torch

    import radik_amd as rk
    task_len = [5000, 70000, 12, 0, 40000, 300000]
    off = np.zeros(len(task_len) + 1, np.int64)
    off[1:] = np.cumsum(task_len)
    K = 2048
    bits = _gen_bits("f32", 1, int(off[-1]), 71)
    vals = torch.from_numpy(bits.view(np.int32).copy()).cuda().view(torch.float32)
    for largest, ascend in ((False, True), (True, False), (True, True)):
        vo, io = rk.topk_radix_select(vals, K, task_len, largest=largest, ascend=ascend)
        torch.cuda.synchronize()
        flags = SORTED | (LARGEST if largest else 0) | (ASCEND if ascend else 0)
        ob, oi = orc.topk(bits, off, None, K=K, flags=flags)
        np.testing.assert_array_equal(io.cpu().numpy().ravel(), oi)
        np.testing.assert_array_equal(vo.cpu().numpy().view(np.uint32).ravel(), ob)


def test_workload_c2_full_size():
    """Config 2 at its full size (512 x 2^17 N(0,1), k = 1024) vs the oracle."""
    import radik_amd.workloads as W
    w = W.c2()
    vals = W.materialize(w)
    bits = from_device_bits(vals, "f32")
    off = w.offsets()
    for exact in (False, True):
        _check("f32", bits, off, w.ks, LARGEST | SORTED | (EXACT if exact else 0))


@pytest.mark.parametrize("dt", ["f32", "bf16"])
def test_workload_c4_scaled(dt):
    """Config 4's shapes (2048 uneven rows, per-row k) with lengths / 16."""
    import radik_amd.workloads as W
    w = W.c4(dt, scale=16)
    vals = W.materialize(w)
    bits = from_device_bits(vals, dt)
    st = _check(dt, bits, w.offsets(), w.ks, LARGEST | SORTED)
    assert st[0] == 0


def test_workload_c3_scaled():
    import radik_amd.workloads as W
    w = W.c3(scale=8, rows=8)
    vals = W.materialize(w)
    bits = from_device_bits(vals, "f32")
    for flags in (LARGEST | SORTED, SORTED | ASCEND):
        for exact in (False, True):
            _check("f32", bits, w.offsets(), w.ks, flags | (EXACT if exact else 0))


def test_repeat_calls_reuse_workspace_deterministic():
    """Same inputs, same workspace, three calls: identical outputs."""
    rng, lens, off = _ragged(81, 40, 1000, 500_000)
    ks = rng.integers(1, 4096, len(lens)).astype(np.int32)
    bits = _gen_bits("f32", 1, int(off[-1]), 82)
    outs = [gpu_topk("f32", bits, off, ks, LARGEST | SORTED)[:2] for _ in range(3)]
    for o in outs[1:]:
        np.testing.assert_array_equal(o[0], outs[0][0])
        np.testing.assert_array_equal(o[1], outs[0][1])


# ---------------------------------------------------------------------------
# Signed zeros, signalling NaNs and subnormals on the streaming filter path.
# Rows are long (>= 70k elements, lengths multiples of 64 so that every row
# starts on a 128-byte line) and their k is small, so they take the sampled
# filter (or, with RK_EXACT, the exact passes and the same filter).  Values
# are built as bit patterns; the smallest-k variants flip every sign bit,
# which mirrors the row.
# ---------------------------------------------------------------------------
SIGN = {"f32": 0x80000000, "bf16": 0x8000, "f16": 0x8000}
SNAN = {"f32": [0x7F800001, 0x7FBFFFFF, 0x7FA00000, 0xFF800001, 0xFFBFFFFF],
        "bf16": [0x7F81, 0x7FBF, 0x7FA0, 0xFF81, 0xFFBF],
        "f16": [0x7C01, 0x7DFF, 0x7D00, 0xFC01, 0xFDFF]}
QNAN = {"f32": [0x7FC00000, 0xFFC00000], "bf16": [0x7FC0, 0xFFC0], "f16": [0x7E00, 0xFE00]}
SUBN_MAX = {"f32": 0x007FFFFF, "bf16": 0x007F, "f16": 0x03FF}


def _bits_of(dt, x):
    """Bit patterns of float32 values x rounded to dt (f16 round to nearest;
    bf16 truncated: the row contents only need to be known bit patterns)."""
    x = np.asarray(x, np.float32)
    if dt == "f32":
        return x.view(np.uint32).copy()
    if dt == "f16":
        return x.astype(np.float16).view(np.uint16).copy()
    return (x.view(np.uint32) >> 16).astype(np.uint16)


def _normal_row(dt, rng, n, scale=1.0):
    return _bits_of(dt, rng.standard_normal(n).astype(np.float32) * np.float32(scale))


def _row_snan_next_to_candidates(dt, rng, n, ncand):
    """N(0,1) row; ncand candidates (50, 50.25, ...) each alone in its
    16-byte vector among signalling NaNs; 500 more vectors of NaNs only."""
    P = 4 if dt == "f32" else 8
    row = _normal_row(dt, rng, n)
    vecs = rng.choice(n // P, ncand + 500, replace=False)
    cand = _bits_of(dt, 50.0 + 0.25 * np.arange(ncand, dtype=np.float32))
    nans = np.array(SNAN[dt] + QNAN[dt], dtype=row.dtype)
    for j, v in enumerate(vecs):
        lane = rng.integers(0, P)
        for c in range(P):
            row[v * P + c] = nans[rng.integers(0, len(SNAN[dt]))] if c != lane or j >= ncand \
                else cand[j]
        if j >= ncand:
            row[v * P + lane] = nans[rng.integers(0, len(nans))]
    return row


def _row_subnormal_kth(dt, rng, n, k):
    """Mostly negative normals; 40 positive normals, 3k distinct positive
    subnormals (some repeated), +0 and -0: the k-th largest is a subnormal.
    Salted with signalling NaNs."""
    row = _normal_row(dt, rng, n)
    row |= np.array(SIGN[dt], dtype=row.dtype)
    pos = rng.choice(n, 40 + 3 * k + 400, replace=False)
    row[pos[:40]] = _bits_of(dt, 1.0 + np.arange(40, dtype=np.float32))
    sub = rng.choice(np.arange(1, SUBN_MAX[dt] + 1), 3 * k, replace=SUBN_MAX[dt] < 3 * k)
    sub[: k // 4] = sub[k // 4: k // 2]  # repeats: ties at subnormal keys
    row[pos[40:40 + 3 * k]] = sub.astype(row.dtype)
    rest = pos[40 + 3 * k:]
    row[rest[:100]] = 0
    row[rest[100:300]] = SIGN[dt]
    row[rest[300:]] = np.array(SNAN[dt], dtype=row.dtype)[rng.integers(0, 5, len(rest) - 300)]
    return row


def _row_zero_boundary(dt, rng, n, k, kth_negative_zero):
    """The row of the +0/-0 boundary: 50 ones, then +0 and -0 filling the
    rest.  With 60 +0 (kth_negative_zero False) the k-th largest (k = 100) is
    +0 and every -0 must stay out; with 30 +0 it is -0 with many ties."""
    dtype = np.uint32 if dt == "f32" else np.uint16
    row = np.full(n, SIGN[dt], dtype)
    pos = rng.choice(n, 50 + 60, replace=False)
    row[pos[:50]] = _bits_of(dt, np.ones(50, np.float32))
    row[pos[50:50 + (30 if kth_negative_zero else 60)]] = 0
    return row


def _mirror(dt, row):
    return row ^ np.array(SIGN[dt], dtype=row.dtype)


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("largest", [True, False])
def test_signed_zero_boundary_long_rows(dt, largest):
    """k-th value +0 in a row of -0 (and the mirror for smallest-k): the
    filter must admit exactly the keys at or above the threshold, not every
    signed zero (they would overflow the row's candidate slots)."""
    rng = np.random.default_rng(91 + DT_CODE[dt])
    n = 1 << 20
    rows = [_row_zero_boundary(dt, rng, n, 100, False), _row_zero_boundary(dt, rng, n, 100, True)]
    if not largest:
        rows = [_mirror(dt, r) for r in rows]
    bits = np.concatenate(rows)
    off = np.array([0, n, 2 * n], np.int64)
    ks = np.array([100, 100], np.int32)
    d = LARGEST if largest else ASCEND
    for flags in (d | SORTED, d):
        for exact in (False, True):
            _check(dt, bits, off, ks, flags | (EXACT if exact else 0))


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("largest", [True, False])
def test_signalling_nan_and_subnormal_long_rows(dt, largest):
    """Signalling-NaN payloads in the same 16-byte vector as the candidates,
    subnormal k-th values (with ties), rows of subnormals only, and the
    +0/-0 boundary, through the sampled filter and the exact passes."""
    rng = np.random.default_rng(71 + 3 * DT_CODE[dt])
    rows = [
        _row_snan_next_to_candidates(dt, rng, 70_016, 300),
        _row_subnormal_kth(dt, rng, 131_072, 100),
        _row_subnormal_kth(dt, rng, 75_008, 1000),
        # subnormals and zeros only (the tiny scale underflows N(0,1))
        _normal_row(dt, rng, 90_048, {"f32": 1e-39, "bf16": 1e-39, "f16": 1e-6}[dt]),
        _row_zero_boundary(dt, rng, 80_000, 100, False),
    ]
    ks = np.array([200, 100, 1000, 1500, 100], np.int32)
    if not largest:
        rows = [_mirror(dt, r) for r in rows]
    lens = np.array([len(r) for r in rows], np.int64)
    off = np.zeros(len(rows) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    bits = np.concatenate(rows)
    # best first, worst first, position order
    flag_sets = ((LARGEST | SORTED, LARGEST | SORTED | ASCEND, LARGEST) if largest else
                 (SORTED | ASCEND, SORTED, 0))
    for flags in flag_sets:
        for exact in (False, True):
            _check(dt, bits, off, ks, flags | (EXACT if exact else 0))


@pytest.mark.parametrize("flags", [LARGEST | SORTED, LARGEST])
@pytest.mark.parametrize("exact", [False, True])
def test_workload_c1_full_size(flags, exact):
    """Config 1 at its full size: one row of 2^20 fp32 U[0,1) on the 2^-24
    grid (seed 1), k = 256, largest; sorted and unsorted, sampled and exact."""
    import radik_amd.workloads as W
    w = W.c1()
    vals = W.materialize(w)
    bits = from_device_bits(vals, "f32")[: w.total]
    st = _check("f32", bits, w.offsets(), w.ks, flags | (EXACT if exact else 0))
    assert st[0] == 0


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("flags", FLAG_SETS)
def test_large_k_chunked_sort(dt, flags):
    """Rows with k above one sort chunk (4096 keys) are sorted in chunks by
    several workgroups and ranked across chunks, on the sampled path and on
    the exact path: k at and around the chunk multiples, duplicate-heavy rows
    (ties broken by position), caller indices, mixed with small-k and
    LDS-direct rows in one call."""
    rng = np.random.default_rng(301 + DT_CODE[dt] * 7 + flags)
    lens = np.array([200_000, 300_001, 150_000, 1 << 20, 90_000, 600_000, 12_000, 70_000],
                    np.int64)
    ks = np.array([4097, 8192, 8193, 16384, 5000, 12_001, 6000, 100], np.int32)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    bits = _gen_bits(dt, 1, int(off[-1]), 303 + flags)
    # ties inside the selection that span chunks: 3000 copies of 2.5 in row 2
    # (about the 0.6 % quantile of N(0,1); the k-th value is near 1.6) and
    # 5000 copies of 2.2 in row 5 (its k-th value is near 2.05)
    for r, v, cnt in ((2, 2.5, 3000), (5, 2.2, 5000)):
        pos = int(off[r]) + rng.choice(int(lens[r]), cnt, replace=False)
        bits[pos] = _bits_of(dt, np.full(cnt, v, np.float32))
    idx_in = rng.integers(-2**31, 2**31 - 1, int(off[-1])).astype(np.int32)
    for exact in (0, EXACT):
        _check(dt, bits, off, ks, flags | exact)
        _check(dt, bits, off, ks, flags | exact, idx_in=idx_in)


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("flags", FLAG_SETS)
def test_bucket_rank_sort_paths(dt, flags):
    """The final sort of a selection by buckets and ranks (2, 4 and 8 keys
    per thread; the 4096-key merge chunks) on selections whose values
    repeat hundreds of times (the buckets must spread copies by position),
    and a selection whose top varying bits are shared by nearly every key
    (a far outlier: the sort falls back to the bitonic network)."""
    rng = np.random.default_rng(901 + DT_CODE[dt] * 11 + flags)
    lens = np.array([400_000, 300_000, 500_000, 1 << 20, 100_000, 80_000, 250_000], np.int64)
    ks = np.array([3000, 4096, 2500, 16384, 1500, 700, 3500], np.int32)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    parts = []
    for r, n in enumerate(lens):
        x = rng.standard_normal(int(n)).astype(np.float32)
        if r in (0, 1, 3, 4, 6):
            x = np.round(x * 4.0).astype(np.float32) / np.float32(4.0)  # values repeat
        if r in (1, 6):
            x[rng.integers(0, n)] = np.float32(3.0e4 if dt == "f16" else 1.0e30)
            x[rng.integers(0, n)] = np.float32(-3.0e4 if dt == "f16" else -1.0e30)
        parts.append(_bits_of(dt, x))
    bits = np.concatenate(parts)
    _check(dt, bits, off, ks, flags)
