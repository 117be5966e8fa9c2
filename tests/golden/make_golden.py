"""Writes tests/golden/topk_golden.json: small top-k known-answer vectors.

Expected outputs come from a brute-force pure-Python implementation that
compares decoded floating-point VALUES (struct 'f' / 'e' decoding, bf16 as
the upper half of an fp32), not the radix key flip the oracle and the HIP
kernels use, so the fixture is an independent check of both:

* order: NaN below every number in both directions; otherwise numeric order
  with -0.0 below +0.0; ties by lower row position;
* the min(len, k) best are selected; sorted output in the requested
  direction (ties: position ascending), unsorted output in position order;
* slots beyond min(len, k) hold value bits 0 and index -1
  (radik/RadixSelect/topk_radixselect.h:312-320).

Run: python tests/golden/make_golden.py   (deterministic; random.Random(2024))
"""

from __future__ import annotations

import json
import math
import os
import random
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "topk_golden.json")

LARGEST, SORTED, ASCEND = 1, 2, 4


def decode(dt: str, bits: int) -> float:
    if dt == "f32":
        return struct.unpack("<f", struct.pack("<I", bits))[0]
    if dt == "f16":
        return struct.unpack("<e", struct.pack("<H", bits))[0]
    return struct.unpack("<f", struct.pack("<I", bits << 16))[0]  # bf16


def encode(dt: str, x: float) -> int:
    if dt == "f32":
        return struct.unpack("<I", struct.pack("<f", x))[0]
    if dt == "f16":
        return struct.unpack("<H", struct.pack("<e", x))[0]
    b = struct.unpack("<I", struct.pack("<f", x))[0]
    return (b + 0x7FFF + ((b >> 16) & 1)) >> 16  # round to nearest even


def rank_tuple(dt: str, bits: int, largest: bool, best_first: bool):
    """Sort tuple (ascending = earlier) for one element, without the index."""
    v = decode(dt, bits)
    if math.isnan(v):
        nan, v, zs = 1, 0.0, 0
    else:
        nan = 0
        zs = 1 if (v == 0.0 and math.copysign(1.0, v) > 0) else 0
    if best_first:
        return (nan, -v, -zs) if largest else (nan, v, zs)
    # worst first: reverse the key order, keep positions ascending
    return (-nan, v, zs) if largest else (-nan, -v, -zs)


def brute_topk(dt: str, row: list, k: int, flags: int, slots: int):
    largest = bool(flags & LARGEST)
    m = max(0, min(len(row), k))
    order = sorted(range(len(row)), key=lambda i: (*rank_tuple(dt, row[i], largest, True), i))
    chosen = order[:m]
    if flags & SORTED:
        best_first = largest != bool(flags & ASCEND)
        chosen.sort(key=lambda i: (*rank_tuple(dt, row[i], largest, best_first), i))
    else:
        chosen.sort()
    vals = [row[i] for i in chosen] + [0] * (slots - m)
    idx = chosen + [-1] * (slots - m)
    return vals, idx


SPECIAL = {
    "f32": [0x00000000, 0x80000000, 0x3F800000, 0xBF800000, 0x7F800000, 0xFF800000,
            0x00000001, 0x80000001, 0x7F7FFFFF, 0xFF7FFFFF],
    "f16": [0x0000, 0x8000, 0x3C00, 0xBC00, 0x7C00, 0xFC00, 0x0001, 0x8001, 0x7BFF, 0xFBFF],
    "bf16": [0x0000, 0x8000, 0x3F80, 0xBF80, 0x7F80, 0xFF80, 0x0001, 0x8001, 0x7F7F, 0xFF7F],
}
NANS = {"f32": [0x7FC00000, 0xFFC00000, 0x7F800001], "f16": [0x7E00, 0xFE00, 0x7C01],
        "bf16": [0x7FC0, 0xFFC0, 0x7F81]}


def random_row(rng: random.Random, dt: str, n: int, style: str) -> list:
    if style == "allequal":
        b = encode(dt, rng.choice([0.0, 1.5, -2.25, 3.0]))
        return [b] * n
    if style == "dups":
        pool = [encode(dt, rng.choice([-1.0, 1.0]) * rng.randint(0, 6) * 0.5) for _ in range(5)]
        return [rng.choice(pool) for _ in range(n)]
    if style == "special":
        out = []
        for _ in range(n):
            r = rng.random()
            if r < 0.25:
                out.append(rng.choice(SPECIAL[dt]))
            elif r < 0.30:
                out.append(rng.choice(NANS[dt]))
            else:
                out.append(encode(dt, rng.gauss(0.0, 3.0)))
        return out
    return [encode(dt, rng.gauss(0.0, 1.0)) for _ in range(n)]  # "normal"


def make_cases():
    rng = random.Random(2024)
    cases = []
    flag_sets = [LARGEST | SORTED, SORTED, LARGEST | SORTED | ASCEND, SORTED | ASCEND, LARGEST, 0]
    for dt in ("f32", "f16", "bf16"):
        for style in ("normal", "dups", "special", "allequal"):
            for flags in flag_sets:
                nrows = rng.randint(1, 4)
                rows, ks = [], []
                for _ in range(nrows):
                    n = rng.choice([0, 1, 2, 3, 5, 8, 17, 33, 64, 100, 129])
                    row = random_row(rng, dt, n, style)
                    k = rng.choice([0, 1, 2, max(1, n - 1), n, n + 3, rng.randint(1, 40)])
                    rows.append(row)
                    ks.append(k)
                slots = [max(0, min(len(r), k)) + rng.choice([0, 0, 2]) for r, k in zip(rows, ks)]
                exp_v, exp_i = [], []
                for row, k, s in zip(rows, ks, slots):
                    v, i = brute_topk(dt, row, k, flags, s)
                    exp_v.append(v)
                    exp_i.append(i)
                cases.append({"dtype": dt, "style": style, "flags": flags, "rows": rows, "ks": ks,
                              "slots": slots, "expect_bits": exp_v, "expect_idx": exp_i})
    return cases


FLIP_KATS = [  # f32 bit pattern -> flipped key (utils.cuh:329-334)
    [0x00000000, 0x80000000], [0x80000000, 0x7FFFFFFF], [0x3F800000, 0xBF800000],
    [0xBF800000, 0x407FFFFF], [0x7F800000, 0xFF800000], [0xFF800000, 0x007FFFFF],
    [0x7FC00000, 0xFFC00000],
]


def main() -> None:
    doc = {"about": "brute-force top-k known answers; see make_golden.py",
           "flip_kats_f32": FLIP_KATS, "cases": make_cases()}
    with open(OUT, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(doc['cases'])} cases")


if __name__ == "__main__":
    main()
