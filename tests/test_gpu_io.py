"""A batch captured to a CSR batch file replays through rk_topk_batched on
the GPU with the flags stored in the file; its result file holds exactly the
oracle's values and indices."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import DT_CODE, NP_BITS, gpu_topk

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("flags", [3, 2, 0])  # largest+sorted, smallest+sorted, smallest unsorted
def test_batch_file_replay(tmp_path, dt, flags):
    import radik_amd as rk
    from oracle import oracle as orc
    rng = np.random.default_rng(11 + flags)
    lens = rng.integers(0, 200_000, 9)
    lens[0] = 0
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    ks = rng.integers(0, 3000, len(lens)).astype(np.int32)
    f32 = rng.standard_normal(int(off[-1])).astype(np.float32)
    if dt == "f32":
        bits = f32.view(np.uint32)
    elif dt == "f16":
        bits = f32.astype(np.float16).view(np.uint16)
    else:
        bits = (f32.view(np.uint32) >> 16).astype(np.uint16)  # truncation is fine for a test
    batch = str(tmp_path / "req.rkb")
    rk.save_batch(batch, DT_CODE[dt], bits, off, ks, flags=flags)

    code, fl, b2, o2, k2 = rk.load_batch(batch)
    assert code == DT_CODE[dt] and fl == flags
    gb, gi, _ = gpu_topk(dt, b2, o2, k2, fl)
    m = np.minimum(o2[1:] - o2[:-1], k2.astype(np.int64))
    out_off = np.zeros(len(k2) + 1, np.int64)
    out_off[1:] = np.cumsum(m)
    res = str(tmp_path / "res.rkr")
    rk.save_results(res, code, out_off, gb, gi)

    _, ro, rv, ri = rk.load_results(res)
    eb, ei = orc.topk(bits.astype(NP_BITS[dt]), off, ks, dt=DT_CODE[dt], flags=flags,
                      out_off=out_off, nthreads=8)
    np.testing.assert_array_equal(ro, out_off)
    np.testing.assert_array_equal(ri, ei)
    np.testing.assert_array_equal(rv, eb)


def test_library_loaded_before_torch():
    """A process that uses the library (here: a batch file) before it imports
    torch still runs the GPU path: the library and torch share one HIP
    runtime.  Fresh interpreter, so the import order is the one written."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys, os, tempfile, numpy as np\n"
        f"sys.path.insert(0, {root!r}); sys.path.insert(0, {os.path.join(root, 'tests')!r})\n"
        "import radik_amd as rk\n"
        "p = os.path.join(tempfile.mkdtemp(), 'b.rkb')\n"
        "bits = np.random.default_rng(1).standard_normal(5000).astype(np.float32).view(np.uint32)\n"
        "rk.save_batch(p, rk.RK_F32, bits, [0, 5000], [10])\n"
        "_, fl, b, o, k = rk.load_batch(p)\n"
        "from helpers import gpu_topk\n"
        "gb, gi, st = gpu_topk('f32', b, o, k, 3)\n"
        "order = np.argsort(-b.view(np.float32), kind='stable')[:10]\n"
        "assert np.array_equal(gi, order), (gi, order)\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
