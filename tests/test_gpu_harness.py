"""The native harness (tests/harness/rk_harness.cpp), the counterpart of the
reference's test_radik.cu command line: batched and per-row-loop calls over
the reference's five value distributions, with and without random task
lengths, each checked bit for bit against the oracle by the harness itself
(--check) and printing the reference's "elapsed: X ms" line."""

from __future__ import annotations

import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(os.path.dirname(__file__), "harness", "_build", "rk_harness")


def _run(*args):
    from oracle import oracle as orc
    orc.lib()  # the harness links the oracle library for --check
    if not os.path.exists(HARNESS):
        pytest.fail(f"{HARNESS} is not built (run __graft_entry__.build())")
    p = subprocess.run([HARNESS, *map(str, args), "--check"], capture_output=True, text=True,
                       timeout=100)
    assert p.returncode == 0, p.stdout + p.stderr
    assert re.search(r"^elapsed: [0-9.]+ ms$", p.stdout, re.M), p.stdout
    assert "check: ok" in p.stdout, p.stdout
    return p.stdout


@pytest.mark.parametrize("dist", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("random_task", [0, 1])
def test_distributions_batched(dist, random_task):
    # BATCH LOG2N K RANDOM_TASK SCALING DISTRIBUTION BATCHED PADDING SORTING
    _run(16, 16, 256, random_task, 1, dist, 1, 0, 1)


@pytest.mark.parametrize("batched", [0, 1])
def test_loop_and_batched_options(batched):
    _run(8, 18, 1024, 1, 1, 0, batched, 0, 1)
    _run(8, 18, 1024, 1, 1, 3, batched, 0, 0, "--largest", "--reps", 3)
    _run(4, 20, 16, 0, 1, 2, batched, 0, 1, "--dtype", "bf16", "--descend")
    _run(4, 12, 4096, 0, 1, 4, batched, 0, 1, "--dtype", "f16", "--largest")
