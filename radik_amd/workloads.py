"""Seeded synthetic workloads of the benchmark and the parity tests.

Shapes follow the five configurations the top-k path is measured on:

* ``c1``: 1 row x 2^20 fp32 U[0,1), k = 256, largest (seed 1);
* ``c2``: 512 rows x 2^17 fp32 N(0,1), k = 1024, largest (seed 2);
* ``c3``: 64 rows x 2^22 fp32, k = 16384: rows 0-31 near-constant
  1 + m*2^-23 (m < 2^13, the top 19 bits identical) with 30 % exact
  duplicates, rows 32-63 Zipf(s = 1.1) ranks over 2^20 values mapped to
  rank / 2^20 (seed 3);
* ``c4``: 2048 rows, len = floor(2^U(10, 24)), k = floor(2^U(0, log2 min(len,
  4096))), N(0,1), fp32, or bf16 / f16 rounded from the same fp32 (seed 4);
* ``c5``: serving requests, len 2^U(12, 22), k 2^U(0, 12), a mixture of
  uniform / normal / Zipf / near-constant rows, Poisson arrivals (seed 5);
* ``adversarial``: equal rows built against the sampled path's sample
  (seed 6), so that every row takes the exact passes.

Row metadata is drawn on the host from a counter-based SplitMix64 stream;
values are generated on the device by ``rk_gen_fill`` (same construction).
``scale`` shrinks every row length (and clamps k) for quick test variants.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from SplitMix64 draws of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = _mix64(np.array([(seed * 0x9E3779B97F4A7C15 ^ stream) & (2**64 - 1)],
                               dtype=np.uint64))[0]
        ctr = base + (np.arange(1, n + 1, dtype=np.uint64) * GAMMA)
    return (_mix64(ctr) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


@dataclass
class Segment:
    """Rows [r0, r1) generated with one generator kind."""
    r0: int
    r1: int
    kind: int
    a: float = 0.0
    b: float = 1.0
    dup: float = 0.0


@dataclass
class Workload:
    name: str
    lens: np.ndarray          # int64 [R]
    ks: np.ndarray            # int32 [R]
    largest: bool
    seed: int
    segments: list = field(default_factory=list)
    dtype: str = "f32"        # "f32" | "bf16" | "f16"
    adversary: str = ""       # "" | "hidden" | "decoy" (see adversarial())

    @property
    def rows(self) -> int:
        return int(self.lens.size)

    @property
    def total(self) -> int:
        return int(self.lens.sum())

    @property
    def max_k(self) -> int:
        return int(self.ks.max()) if self.ks.size else 0

    def offsets(self) -> np.ndarray:
        off = np.zeros(self.rows + 1, dtype=np.int64)
        np.cumsum(self.lens, out=off[1:])
        return off

    def out_offsets(self) -> np.ndarray:
        m = np.minimum(self.lens, self.ks.astype(np.int64))
        off = np.zeros(self.rows + 1, dtype=np.int64)
        np.cumsum(m, out=off[1:])
        return off

    def algorithmic_bytes(self) -> int:
        """One read of every value + (value, int32 index) written per selected
        element + 16 B of row metadata (offset, length, k)."""
        es = 4 if self.dtype == "f32" else 2
        m = np.minimum(self.lens, self.ks.astype(np.int64))
        return int(self.total * es + int(m.sum()) * (es + 4) + 16 * self.rows)


# generator kinds (radik_amd_harness.h)
UNIFORM, NORMAL, ZIPF, NEARCONST, RANGE, CONST = range(6)


def c1(scale: int = 1) -> Workload:
    n = max(1024, (1 << 20) // scale)
    return Workload("c1_plain_single", np.array([n], np.int64), np.array([min(256, n)], np.int32),
                    True, 1, [Segment(0, 1, UNIFORM)])


def c2(scale: int = 1, rows: int = 512) -> Workload:
    n = max(1024, (1 << 17) // scale)
    return Workload("c2_plain_batched", np.full(rows, n, np.int64),
                    np.full(rows, min(1024, n), np.int32), True, 2,
                    [Segment(0, rows, NORMAL, 0.0, 1.0)])


def c3(scale: int = 1, rows: int = 64, largest: bool = True) -> Workload:
    n = max(1 << 15, (1 << 22) // scale)
    h = rows // 2
    k = min(16384, n // 4)
    return Workload("c3_skewed", np.full(rows, n, np.int64), np.full(rows, k, np.int32), largest,
                    3, [Segment(0, h, NEARCONST, dup=0.3), Segment(h, rows, ZIPF, 1.1, 2.0 ** 20)])


def c4(dtype: str = "f32", scale: int = 1, rows: int = 2048) -> Workload:
    u = uniform01(4, 1, rows)
    lo, hi = 10, 24
    lens = np.floor(2.0 ** (lo + (hi - lo) * u)).astype(np.int64)
    lens = np.maximum(lens // scale, 1)
    cap = np.log2(np.minimum(lens, 4096).astype(np.float64))
    ks = np.floor(2.0 ** (uniform01(4, 2, rows) * cap)).astype(np.int32)
    ks = np.maximum(ks, 1)
    return Workload(f"c4_batched_uneven_{dtype}", lens, ks, True, 4,
                    [Segment(0, rows, NORMAL, 0.0, 1.0)], dtype)


def sample_positions(n: int, sample_groups: int = 1024, group_elems: int = 16, pack: int = 4,
                     a0: int = 0, full: bool = False) -> np.ndarray:
    """Row positions the sampled path reads for a row of n elements: runs of
    group_elems keys read as 16-byte vectors (pack elements each), run g
    starting at a0 + floor(g * (n - a0 - group_elems) / (ng - 1)) rounded down
    to a multiple of pack, with a0 the row's first element on a 16-byte
    boundary and ng the number of runs the library derives from sample_groups
    and n (at most 16384 keys, about 1/64 of the row unless `full` (rows with
    k above the LDS select take the full sample), at least 64 runs).
    Used to build inputs that defeat the sample; the sampled path itself is
    exact whatever the sample sees."""
    ge = group_elems
    keys = min(16384, max(sample_groups, 64) * ge)
    ng = sample_groups
    if not full and ng * ge * 64 > n:
        ng = n // (64 * ge)
    ng = max(ng, min(n // (2 * ge), 64))
    ng = max(min(ng, keys // ge), 1)
    step = (n - a0 - ge) / (ng - 1) if ng > 1 else 0.0
    starts = a0 + (np.arange(ng) * step).astype(np.int64) // pack * pack
    return (starts[:, None] + np.arange(ge)[None, :]).ravel()


def adversarial(kind: str, rows: int = 64, n: int = 1 << 22, k: int = 1024, dtype: str = "f32",
                largest: bool = True) -> Workload:
    """Rows built against the sample of the sampled path (seed 6), N(0,1)
    values shifted by 8 in the direction of the selection at chosen positions:

    * ``hidden``: every position except the sampled ones (for 512 and 1024
      runs of 16 keys) is shifted: the sample sees none of the best keys, its
      threshold admits nearly the whole row and the candidates overflow;
    * ``decoy``: only the sampled positions (1024 runs of 16 keys, the default
      tunables) are shifted: the sample sees only the best keys, and fewer
      than k keys of the row reach its threshold.

    Every row falls back to the exact passes; the output is still exact."""
    if kind not in ("hidden", "decoy"):
        raise ValueError(f"unknown adversary {kind!r}")
    return Workload(f"adversarial_{kind}_{dtype}", np.full(rows, n, np.int64),
                    np.full(rows, min(k, n), np.int32), largest, 6,
                    [Segment(0, rows, NORMAL, 0.0, 1.0)], dtype, kind)


def _apply_adversary(w: Workload, vals, off) -> None:
    import torch

    shift = 8.0 if w.largest else -8.0
    pack = 4 if w.dtype == "f32" else 8  # elements per 16-byte vector
    for n in np.unique(w.lens):
        # rows of equal length n start on 16-byte boundaries here (16 | n * size)
        assert (int(n) * (4 if w.dtype == "f32" else 2)) % 16 == 0
        if w.adversary == "decoy":
            pos = sample_positions(int(n), 1024, 16, pack)
        else:
            seen = np.union1d(sample_positions(int(n), 1024, 16, pack),
                              sample_positions(int(n), 512, 16, pack))
            pos = np.setdiff1d(np.arange(int(n)), seen)
        d_pos = torch.from_numpy(pos).to(vals.device)
        for r in np.nonzero(w.lens == n)[0]:
            row = vals[int(off[r]):int(off[r + 1])]
            row.index_add_(0, d_pos, torch.full((len(pos),), shift, dtype=vals.dtype,
                                                 device=vals.device))


def c5_requests(n: int = 100_000, scale: int = 1):
    """(lens, ks, kinds) of n serving requests (mixture of four families)."""
    lens = np.floor(2.0 ** (12 + 10 * uniform01(5, 1, n))).astype(np.int64)
    lens = np.maximum(lens // scale, 1)
    ks = np.floor(2.0 ** (12 * uniform01(5, 2, n))).astype(np.int32)
    ks = np.clip(ks, 1, None)
    kinds = np.floor(4 * uniform01(5, 3, n)).astype(np.int32)  # 0 U, 1 N, 2 Zipf, 3 near-const
    return lens, ks, kinds


def poisson_arrivals(n: int, rate_per_s: float, seed: int = 5) -> np.ndarray:
    """Arrival times (s) of a Poisson process with the given rate."""
    u = uniform01(seed, 4, n)
    gaps = -np.log1p(-u) / rate_per_s
    return np.cumsum(gaps)


MIX_PARAMS = {0: (UNIFORM, 0.0, 1.0, 0.0), 1: (NORMAL, 0.0, 1.0, 0.0),
              2: (ZIPF, 1.1, 2.0 ** 20, 0.0), 3: (NEARCONST, 0.0, 0.0, 0.3)}


def materialize_requests(lens, kinds, seed: int = 5, dtype: str = "f32", device="cuda"):
    """One device arena holding every request's row (request i at
    [off[i], off[i+1]), family MIX_PARAMS[kinds[i]], generator stream i);
    returns (values, off)."""
    import torch

    import radik_amd as rk

    lens = np.asarray(lens, np.int64)
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    tdt = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[dtype]
    vals = torch.empty(max(int(off[-1]), 1), dtype=tdt, device=device)
    for i, kind in enumerate(np.asarray(kinds)):
        lo, hi = int(off[i]), int(off[i + 1])
        if hi > lo:
            g, a, b, dup = MIX_PARAMS[int(kind)]
            rk.gen_fill(vals[lo:hi], g, seed, i, a=a, b=b, dup_frac=dup, row_len=hi - lo)
    return vals, off


def c5_bodies(per_kind: int = 256, body_len: int = 1 << 22, seed: int = 5, device="cuda"):
    """Resident request bodies of the serving workload: per_kind fp32 rows of
    body_len for each of the four families; body (kind, j) is row
    kind * per_kind + j.  A request of length n reads the first n elements of
    a body of its family."""
    import torch

    import radik_amd as rk

    bodies = torch.empty((4 * per_kind, body_len), dtype=torch.float32, device=device)
    for kind in range(4):
        g, a, b, dup = MIX_PARAMS[kind]
        for j in range(per_kind):
            r = kind * per_kind + j
            rk.gen_fill(bodies[r], g, seed, 1000 + r, a=a, b=b, dup_frac=dup, row_len=body_len)
    return bodies


def materialize(w: Workload, device="cuda"):
    """Device tensor of the workload's values (generated on the device)."""
    import torch

    import radik_amd as rk

    tdt = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[w.dtype]
    off = w.offsets()
    vals = torch.empty(max(w.total, 1), dtype=tdt, device=device)
    if w.dtype != "f32":
        # 16-bit runs round the same fp32 values (round to nearest even); an
        # adversary is placed after rounding, on the 16-bit sample's grid
        f32 = materialize(Workload(w.name, w.lens, w.ks, w.largest, w.seed, w.segments, "f32"),
                          device)
        vals.copy_(f32[: vals.numel()])
        del f32
        if w.adversary:
            _apply_adversary(w, vals, off)
        return vals
    for seg in w.segments:
        for r in range(seg.r0, seg.r1):
            lo, hi = int(off[r]), int(off[r + 1])
            if hi > lo:
                rk.gen_fill(vals[lo:hi], seg.kind, w.seed, r, a=seg.a, b=seg.b,
                            dup_frac=seg.dup, row_len=hi - lo)
    if w.adversary:
        _apply_adversary(w, vals, off)
    return vals
