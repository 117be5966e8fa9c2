"""bench.py's contract pieces that need no GPU: the refusal of more than one
GPU (exit code 2, before any device call), the C4 workload definition, and
the nearest-rank percentile the serving statistics use."""

from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_c4_workload_definition():
    sys.path.insert(0, ROOT)
    import radik_amd.workloads as W
    for dt in ("f32", "bf16"):
        w = W.c4(dt)
        assert w.rows == 2048 and w.dtype == dt
        assert w.total == int(w.lens.sum()) == 3_391_571_209
        assert w.lens.min() >= 1 << 10 and w.lens.max() < 1 << 24
        assert np.all(w.ks >= 1) and np.all(w.ks <= np.minimum(w.lens, 4096))
    # the same seed gives the same rows for both dtypes
    a, b = W.c4("f32"), W.c4("bf16")
    assert np.array_equal(a.lens, b.lens) and np.array_equal(a.ks, b.ks)


def test_percentile_is_nearest_rank():
    sys.path.insert(0, ROOT)
    import bench
    x = np.arange(1, 101, dtype=np.float64)  # 1..100
    assert bench.percentile(x, 0.50) == 50.0
    assert bench.percentile(x, 0.99) == 99.0
    assert bench.percentile(x, 1.00) == 100.0
    assert bench.percentile([7.0], 0.99) == 7.0
    assert bench.percentile([], 0.5) == 0.0


def test_sample_positions_cover_evenly_spaced_runs():
    """The adversarial workloads' picture of the sample: ng runs of 16 keys on
    the row's 16-byte vector grid (pack elements per vector, from a0), the
    first at a0, within pack of g * (n - a0 - 16) / (ng - 1) + a0, the last
    inside the row, at most 16384 keys and about 1/64 of a row."""
    sys.path.insert(0, ROOT)
    import radik_amd.workloads as W
    for n, sg, ng in ((1 << 20, 1024, 1024), (1 << 22, 512, 512), (1 << 22, 2048, 1024),
                      (1 << 16, 1024, 64), (5000, 1024, 64)):
        for pack, a0 in ((4, 0), (8, 0), (4, 3), (8, 5)):
            p = W.sample_positions(n, sg, 16, pack, a0)
            assert len(p) == ng * 16
            runs = p.reshape(ng, 16)
            assert (np.diff(runs, axis=1) == 1).all()
            assert runs[0, 0] == a0 and runs[-1, -1] <= n - 1
            assert ((runs[:, 0] - a0) % pack == 0).all()
            ideal = a0 + np.arange(ng) * ((n - a0 - 16) / (ng - 1))
            assert (runs[:, 0] <= ideal + 1e-6).all() and (runs[:, 0] > ideal - pack).all()
    w = W.adversarial("hidden", rows=2, n=1 << 20)
    assert w.adversary == "hidden" and w.rows == 2 and int(w.ks[0]) == 1024
