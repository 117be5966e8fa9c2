"""Benchmark of the batched radix top-k on one MI355X.

Default workload: batched uneven rows (2048 rows, len = floor(2^U(10,24)),
k = floor(2^U(0, log2 min(len, 4096))), N(0,1) fp32, largest, sorted),
one rk_topk_batched call per step with inputs resident in HBM.

Prints one JSON line with:
  value        whole-job keys/s (sum of row lengths / time per step)
  roofline     the dominant kernel (the streaming filter): algorithmic bytes
               per launch / its average duration, measured with HIP events
               recorded around it inside every timed step
  cpu_baseline the multithreaded CPU oracle (std::nth_element + std::sort)
               on a bounded sample of the same rows, on this host's cores
  plain        the same rows with one RadiK multi-pass call per row
  serving      the serving loop on the mixed request stream (100k requests,
               Poisson arrivals): p50/p99 latency and throughput of the
               continuous-batching server and of per-request plain calls at
               the same offered rate
Progress lines go to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "top-k input keys/s + % HBM roofline (batched vs plain); serving p50/p99 latency"
HBM_PEAK_GBS = 8000.0
REPS = 20  # individually timed graph replays (median and spread)


def log(msg: str) -> None:
    sys.stderr.write(f"[bench {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.stderr.flush()


PROFILE_KEY = {"c4_batched_uneven_f32": "c4_f32", "c4_batched_uneven_bf16": "c4_bf16",
               "c4_batched_uneven_f16": "c4_f16"}


def profiled_traffic(workload: str) -> dict:
    """HBM bytes per k_filter launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes summarized in profiles/k_filter_<workload>.txt, with what that
    figure rests on: the profiled run's tunables and the scale factors that
    turn the counters into bytes (each calibrated on a kernel of known size:
    reads on the read probe, writes on the input generator), so the figure is
    relative to those kernels.  This run's own tunables may differ (the
    autotuner's winner varies from run to run within ~1 % of the step)."""
    path = os.path.join(ROOT, "profiles", f"k_filter_{workload}.txt")
    out = {"traffic": None}
    if not os.path.exists(path):
        return out
    src = {"profile": os.path.relpath(path, ROOT)}
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                out["traffic"] = float(line.split(":", 1)[1])
            elif line.startswith("# tunables"):
                src["profiled_tunables"] = line.split(":", 1)[1].strip()
            elif "read scale" in line and "=" in line:
                src["read_scale"] = float(line.rsplit("=", 1)[1])
            elif "write scale" in line and "=" in line:
                src["write_scale"] = float(line.rsplit("=", 1)[1])
            elif line.startswith("traffic_over_algorithmic:"):
                src["traffic_over_algorithmic"] = float(line.split(":", 1)[1])
    src["calibration"] = ("FETCH_SIZE scaled by the read probe's bytes / its FETCH_SIZE, "
                          "WRITE_SIZE by the generator's bytes / its WRITE_SIZE")
    out["traffic_source"] = src
    return out


def make_workload(name: str):
    import radik_amd.workloads as W
    table = {
        "c4_f32": lambda: W.c4("f32"),
        "c4_bf16": lambda: W.c4("bf16"),
        "c4_f16": lambda: W.c4("f16"),
        "c2": lambda: W.c2(),
        "c1": lambda: W.c1(),
        "c3": lambda: W.c3(),
    }
    return table[name]()


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c4_f32",
                    choices=["c4_f32", "c4_bf16", "c4_f16", "c2", "c1", "c3", "knn"],
                    help="knn: the fused k-nearest-neighbour select against its unfused "
                         "form (prints its own JSON object)")
    ap.add_argument("--knn-shape", default="1024,1048576,128,100",
                    help="knn workload: queries,database vectors,dim,k")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default tunables instead of the autotuned ones")
    ap.add_argument("--cpu-sample-keys", type=float, default=4.0e9,
                    help="bound on the keys in the CPU baseline's sample")
    ap.add_argument("--serve-requests", type=int, default=100_000,
                    help="requests in the serving run (0: skip it)")
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="offered rate as a fraction of the plain server's capacity")
    ap.add_argument("--serve-hi-load", type=float, default=0.5,
                    help="second offered rate (batched server only), as a fraction of the "
                         "batched server's capacity")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving measurement (prints its JSON object)")
    ap.add_argument("--no-bf16", action="store_true",
                    help="default workload: skip the bf16 run of the same batch")
    ap.add_argument("--no-sweeps", dest="sweeps", action="store_false",
                    help="skip the single-row and quantile sweeps")
    ap.add_argument("--sweeps-only", action="store_true",
                    help="run only the sweeps (prints their JSON object)")
    ap.add_argument("--no-fallback", dest="fallback", action="store_false",
                    help="skip the fallback-cost measurement (sampler-defeating rows)")
    ap.add_argument("--fallback-only", action="store_true",
                    help="run only the fallback-cost measurement (prints its JSON object)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; rows are independent, so "
                          "a batch would shard across GPUs by rows with no exchange",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W

    torch.cuda.set_device(0)
    if args.serve_only:
        print(json.dumps({"serving": serving(args)}), flush=True)
        return 0
    if args.workload == "knn":
        print(json.dumps({"knn": knn_bench(args)}), flush=True)
        return 0
    if args.sweeps_only:
        print(json.dumps({"sweeps": sweeps(args)}), flush=True)
        return 0
    if args.fallback_only:
        print(json.dumps({"fallback_cost": fallback_cost(args)}), flush=True)
        return 0
    w = make_workload(args.workload)
    result = measure(w, args, primary=True)
    if result.get("bit_exact") is False:
        print(json.dumps(result), flush=True)
        return 1
    torch.cuda.empty_cache()
    if args.workload == "c4_f32" and not args.no_bf16:
        # the same batch in bf16 (rounded from the same fp32 values), on the
        # same clock: its own roofline, plain path, CPU baseline and check
        extra = measure(make_workload("c4_bf16"), args, primary=False)
        result["c4_bf16"] = extra
        torch.cuda.empty_cache()
        if extra.get("bit_exact") is False:
            print(json.dumps(result), flush=True)
            return 1
    if args.sweeps:
        result["sweeps"] = sweeps(args)
        torch.cuda.empty_cache()
    if args.fallback:
        result["fallback_cost"] = fallback_cost(args)
        torch.cuda.empty_cache()

    if args.serve_requests > 0:
        result["serving"] = serving(args)

    print(json.dumps(result), flush=True)
    return 0


def measure(w, args, primary: bool) -> dict:
    """One batched workload at full size: autotune, eager and graph-replay
    timing, the filter kernel's roofline, the read probe, the plain per-row
    path, the CPU baseline and the bit-exact check of the timed output.
    ``primary``: the fields of the top-level bench line (metric, value, ...)."""
    import numpy as np
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W

    es = 4 if w.dtype == "f32" else 2
    dt = {"f32": rk.RK_F32, "bf16": rk.RK_BF16, "f16": rk.RK_F16}[w.dtype]
    log(f"workload {w.name}: {w.rows} rows, {w.total} keys, max k {w.max_k}")
    vals = W.materialize(w)
    off = w.offsets()
    oo = w.out_offsets()
    d_off = torch.from_numpy(off).cuda()
    d_k = torch.from_numpy(w.ks).cuda()
    d_oo = torch.from_numpy(oo).cuda()
    n_out = int(oo[-1])
    out_v = torch.empty(n_out, dtype=vals.dtype, device="cuda")
    out_i = torch.empty(n_out, dtype=torch.int32, device="cuda")
    flags = rk.RK_LARGEST | rk.RK_SORTED
    ws_bytes = rk.workspace_size(dt, w.rows, w.total, w.max_k, flags)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    L = rk.lib()
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    log(f"inputs resident: {vals.numel() * es / 1e9:.2f} GB values, workspace {ws_bytes / 1e6:.1f} MB")

    def step():
        rc = L.rk_topk_batched(dt, vals.data_ptr(), d_off.data_ptr(), d_k.data_ptr(), w.rows,
                               w.total, w.max_k, flags, None, out_v.data_ptr(), out_i.data_ptr(),
                               d_oo.data_ptr(), ws.data_ptr(), ws_bytes, sp)
        if rc != 0:
            raise rk.RadikError(rc, "rk_topk_batched")

    tuned = None
    if not args.no_autotune:
        # The autotuner times its candidate tunables on these buffers (untimed
        # here) and caches the winner for this shape bucket; every later call
        # in the bucket, eager or captured, uses it.
        a0 = time.perf_counter()
        tuned, tuned_ms = rk.autotune_batched(vals, d_off, d_k, out_v, out_i, d_oo, ws,
                                              total_len=w.total, max_k_=w.max_k, reps=3,
                                              stream=stream)
        log(f"autotuned in {time.perf_counter() - a0:.1f} s: {tuned} ({tuned_ms:.3f} ms)")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    st = rk.status(ws)
    if st[0] != 0:
        raise RuntimeError(f"device status {st}")
    log(f"warmup done ({args.warmup} steps); fallback rows {st[1]}, ordered rows {st[2]}")

    rk.timing_read()  # drop anything recorded before
    rk.timing_enable(True)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(args.steps):
        step()
    t1.record(stream)
    torch.cuda.synchronize()
    rk.timing_enable(False)
    elapsed_ms = t0.elapsed_time(t1)
    filt_ms, filt_n = rk.timing_read()
    ms_step = elapsed_ms / args.steps
    keys_s = w.total / (ms_step / 1e3)
    log(f"batched: {ms_step:.3f} ms/step, {keys_s / 1e9:.2f} Gkeys/s")

    # The same call captured once in a hipGraph and replayed: the batched
    # path's launch form for a fixed set of buffers (one graph launch per step).
    gstream = torch.cuda.Stream()
    torch.cuda.synchronize()
    with rk.Graph.capture(gstream) as graph:
        rc = L.rk_topk_batched(dt, vals.data_ptr(), d_off.data_ptr(), d_k.data_ptr(), w.rows,
                               w.total, w.max_k, flags, None, out_v.data_ptr(),
                               out_i.data_ptr(), d_oo.data_ptr(), ws.data_ptr(), ws_bytes,
                               gstream.cuda_stream)
        if rc != 0:
            raise rk.RadikError(rc, "rk_topk_batched (capture)")
    for _ in range(args.warmup):
        graph.launch(gstream)
    gstream.synchronize()
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    g0.record(gstream)
    for _ in range(args.steps):
        graph.launch(gstream)
    g1.record(gstream)
    gstream.synchronize()
    graph_ms = g0.elapsed_time(g1) / args.steps
    # Per-replay times as well (a separate pass, so the bracketed steps above
    # carry no per-step events): the median and spread of 20 replays.
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(REPS + 1)]
    ev[0].record(gstream)
    for i in range(REPS):
        graph.launch(gstream)
        ev[i + 1].record(gstream)
    gstream.synchronize()
    reps_ms = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(REPS)])
    graph.close()
    st = rk.status(ws)
    if st[0] != 0:
        raise RuntimeError(f"device status after graph replays {st}")
    # The timed calls' output (the last replay's), kept for the bit-exact
    # check against the oracle in the CPU-baseline leg.
    timed_out = (out_v.view(torch.int32 if es == 4 else torch.int16).cpu().numpy(),
                 out_i.cpu().numpy())
    log(f"graph replay: {graph_ms:.3f} ms/step ({w.total / (graph_ms / 1e3) / 1e9:.2f} Gkeys/s)")
    eager_ms = ms_step
    ms_step = min(ms_step, graph_ms)
    keys_s = w.total / (ms_step / 1e3)

    # Dominant kernel: the streaming filter.  Per launch it reads every key of
    # the rows longer than the LDS-direct limit once: sizeof(T) bytes per key.
    # Rows the filter streams: longer than the LDS-direct limit of this call
    # (next power of two >= max(4096, max_k), as the library sizes it) and
    # longer than their k (shorter rows are copied whole).
    small_max = 4096
    while small_max < w.max_k:
        small_max *= 2
    streamed = (w.lens > small_max) & (w.lens > w.ks)
    long_keys = int(w.lens[streamed].sum())
    filt_bytes = long_keys * es
    filt_avg_ms = filt_ms / max(filt_n, 1)
    achieved = filt_bytes / (filt_avg_ms / 1e3) / 1e9
    alg_bytes = w.algorithmic_bytes()
    pipeline_gbs = alg_bytes / (ms_step / 1e3) / 1e9

    result = {
        "metric": METRIC,
        "value": keys_s,
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": w.dtype,
        "data": "synthetic: seeded N(0,1) generated on the device (counter-based SplitMix64)",
        "config": {"workload": w.name, "rows": w.rows, "total_keys": w.total,
                   "len": "floor(2^U(10,24))", "k": "floor(2^U(0, log2 min(len,4096)))",
                   "largest": True, "sorted": True, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_filter", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     **profiled_traffic(PROFILE_KEY.get(w.name, w.name)),
                     "algorithmic_bytes_per_launch": filt_bytes,
                     "avg_launch_ms": filt_avg_ms, "launches_timed": filt_n},
        "pipeline": {"algorithmic_bytes": alg_bytes, "achieved_gbs": pipeline_gbs,
                     "roofline_frac": pipeline_gbs / HBM_PEAK_GBS,
                     "eager_ms_per_step": eager_ms, "graph_ms_per_step": graph_ms,
                     "graph_replay_ms": {"reps": REPS, "median": float(np.median(reps_ms)),
                                         "min": float(reps_ms.min()),
                                         "max": float(reps_ms.max())},
                     "tunables": tuned or "defaults",
                     "fallback_rows": st[1], "ordered_rows": st[2]},
    }

    # Achievable read bandwidth of this box: the streaming-read probe over
    # the same resident values, the fastest of 2, 3 and 4 workgroups per CU.
    probe_gbs = 0.0
    for bpc in (2, 3, 4):
        rk.probe_read(vals, bpc)
        torch.cuda.synchronize()
        q0 = torch.cuda.Event(enable_timing=True)
        q1 = torch.cuda.Event(enable_timing=True)
        q0.record(stream)
        for _ in range(5):
            rk.probe_read(vals, bpc)
        q1.record(stream)
        torch.cuda.synchronize()
        probe_ms = q0.elapsed_time(q1) / 5
        probe_gbs = max(probe_gbs, vals.numel() * es / (probe_ms / 1e3) / 1e9)
    result["roofline"]["achievable_peak"] = probe_gbs
    result["roofline"]["frac_of_achievable"] = achieved / probe_gbs
    log(f"stream-read probe: {probe_gbs:.0f} GB/s; filter {achieved:.0f} GB/s")

    if not args.no_plain and args.plain_steps > 0:
        # One exact multi-pass call per row (RadiK's per-query use of
        # topKRadixSelectL with taskNum = 1, radik/test/test_radik.cu:164-183).
        pflags = flags | rk.RK_EXACT
        p_ws = rk.workspace_size(dt, 1, int(w.lens.max()), w.max_k, pflags)
        pws = torch.empty(p_ws, dtype=torch.uint8, device="cuda")
        es_v = vals.element_size()

        def plain_step():
            for r in range(w.rows):
                rc = L.rk_topk_batched(dt, vals.data_ptr(), d_off.data_ptr() + 8 * r,
                                       d_k.data_ptr() + 4 * r, 1, int(w.lens[r]), int(w.ks[r]),
                                       pflags, None, out_v.data_ptr(), out_i.data_ptr(),
                                       d_oo.data_ptr() + 8 * r, pws.data_ptr(), p_ws, sp)
                if rc != 0:
                    raise rk.RadikError(rc, "plain rk_topk_batched")
        del es_v
        plain_step()
        torch.cuda.synchronize()
        p0 = torch.cuda.Event(enable_timing=True)
        p1 = torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        for _ in range(args.plain_steps):
            plain_step()
        p1.record(stream)
        torch.cuda.synchronize()
        p_ms = p0.elapsed_time(p1) / args.plain_steps
        result["plain"] = {"ms_per_step": p_ms, "keys_per_s": w.total / (p_ms / 1e3),
                           "batched_speedup": p_ms / ms_step,
                           "what": "one RK_EXACT (multi-pass) call per row, same stream"}
        log(f"plain per-row: {p_ms:.2f} ms/step ({p_ms / ms_step:.1f}x slower)")

    if not args.no_cpu:
        from oracle import oracle as orc
        # bounded sample: every s-th row, at most cpu_sample_keys keys
        stride = max(1, int(np.ceil(w.total / args.cpu_sample_keys)))
        rows = np.arange(0, w.rows, stride)
        lens = w.lens[rows]
        s_off = np.zeros(len(rows) + 1, np.int64)
        s_off[1:] = np.cumsum(lens)
        host = np.empty(int(s_off[-1]), dtype=np.uint32 if es == 4 else np.uint16)
        iview = vals.view(torch.int32 if es == 4 else torch.int16)
        for j, r in enumerate(rows):
            host[s_off[j]:s_off[j + 1]] = iview[off[r]:off[r + 1]].cpu().numpy().view(host.dtype)
        nth = orc.usable_threads()
        c0 = time.perf_counter()
        eb, ei = orc.topk(host, s_off, w.ks[rows], dt=dt, flags=orc.LARGEST | orc.SORTED,
                          nthreads=nth)
        c_s = time.perf_counter() - c0
        # The timed GPU output of the sampled rows must equal the oracle's.
        sel = np.concatenate([np.arange(oo[r], oo[r + 1]) for r in rows])
        gv = timed_out[0][sel].view(eb.dtype)
        exact = bool(np.array_equal(gv, eb) and np.array_equal(timed_out[1][sel], ei))
        result["bit_exact"] = exact
        result["bit_exact_checked"] = {"rows": int(len(rows)), "outputs": int(len(sel)),
                                       "against": "oracle (CPU restatement), values and indices"}
        if not exact:
            log("timed output differs from the oracle")
            return result
        result["cpu_baseline"] = {
            "value": float(s_off[-1]) / c_s, "unit": "keys/s", "cores": nth, "kind": "port",
            "cpu_model": cpu_model(), "host_threads": orc.hardware_threads(),
            "threads_rule": "every CPU of this process's affinity mask within the cgroup quota",
            "sample": ("the whole workload" if stride == 1 else
                       f"every {stride}th row of the workload")
                      + f" ({len(rows)} rows, {int(s_off[-1])} keys), std::nth_element + "
                      f"std::sort per row, {nth} threads, {c_s:.2f} s"}
        log(f"cpu baseline: {float(s_off[-1]) / c_s / 1e9:.3f} Gkeys/s on {nth} threads")
    if not primary:
        for key in ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling",
                    "vs_baseline"):
            result.pop(key, None)
    del vals, ws, out_v, out_i
    return result


def sweeps(args) -> dict:
    """The reference's own benchmark sweeps, timed on this path (values
    U[0,1) fp32 from a seed, generated on the device; every point is one call
    captured in a hipGraph and replayed, default tunables):

      single_query  one row of N = 2^21 .. 2^29 keys, K = 2^4 .. 2^12, largest,
                    sorted (eval/1-simple-topk.py:18-27); the sampled path and
                    the exact multi-pass path (RK_EXACT), and the streaming
                    filter's own bandwidth (HIP events around it);
      quantile      16 rows of N = 2^16 .. 2^22 keys, K = 512, N/100, N/4 and
                    N/2, unsorted (eval/ex1-median.py:11-26).

    Roofline: algorithmic bytes (one read of every key, K x (value + index)
    written, 16 B per row) / time / 8 TB/s."""
    import numpy as np
    import torch

    import radik_amd as rk

    L = rk.lib()
    gstream = torch.cuda.Stream()
    sp = gstream.cuda_stream
    reps = 10

    def replay_ms(call):
        torch.cuda.synchronize()
        with rk.Graph.capture(gstream) as g:
            call()
        for _ in range(2):
            g.launch(gstream)
        gstream.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(gstream)
        for _ in range(reps):
            g.launch(gstream)
        e1.record(gstream)
        gstream.synchronize()
        g.close()
        return e0.elapsed_time(e1) / reps

    nmax, kmax = 1 << 29, 1 << 12
    vals = torch.empty(nmax, dtype=torch.float32, device="cuda")
    rk.gen_fill(vals, rk.GEN_UNIFORM, 11, 0)
    flags = rk.RK_LARGEST | rk.RK_SORTED
    ws_bytes = max(rk.workspace_size(rk.RK_F32, 1, nmax, kmax, flags),
                   rk.workspace_size(rk.RK_F32, 1, nmax, kmax, flags | rk.RK_EXACT))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    ov = torch.empty(kmax, dtype=torch.float32, device="cuda")
    oi = torch.empty(kmax, dtype=torch.int32, device="cuda")
    t0 = time.perf_counter()
    single = []
    for ln in range(21, 30):
        n = 1 << ln
        d_off = torch.tensor([0, n], dtype=torch.int64, device="cuda")
        for ko in range(4, 13):
            k = 1 << ko
            d_k = torch.tensor([k], dtype=torch.int32, device="cuda")
            d_oo = torch.tensor([0, k], dtype=torch.int64, device="cuda")

            def call(fl):
                rc = L.rk_topk_batched(rk.RK_F32, vals.data_ptr(), d_off.data_ptr(),
                                       d_k.data_ptr(), 1, n, k, fl, None, ov.data_ptr(),
                                       oi.data_ptr(), d_oo.data_ptr(), ws.data_ptr(), ws_bytes, sp)
                if rc != 0:
                    raise rk.RadikError(rc, "rk_topk_batched (sweep)")

            ms = replay_ms(lambda: call(flags))
            st = rk.status(ws, gstream)
            # the filter alone, eager, with HIP events around it
            rk.timing_read()
            rk.timing_enable(True)
            for _ in range(3):
                call(flags)
            gstream.synchronize()
            rk.timing_enable(False)
            f_ms, f_n = rk.timing_read()
            ems = replay_ms(lambda: call(flags | rk.RK_EXACT))
            est = rk.status(ws, gstream)
            if st[0] or est[0]:
                raise RuntimeError(f"sweep N=2^{ln} K={k}: device status {st} / {est}")
            alg = n * 4 + k * 8 + 16
            single.append([n, k, round(ms, 5), n / (ms / 1e3), round(alg / (ms / 1e3) / 1e9 /
                           HBM_PEAK_GBS, 4), round(ems, 5),
                           round(n * 4 / (f_ms / max(f_n, 1) / 1e3) / 1e9, 1) if f_n else None])
    big = [r for r in single if r[0] == nmax]
    log(f"sweep single_query: {len(single)} points in {time.perf_counter() - t0:.1f} s; "
        f"N=2^29: {min(r[2] for r in big):.3f}-{max(r[2] for r in big):.3f} ms, "
        f"filter {max(r[6] for r in big):.0f} GB/s")
    del ws
    # quantile: 16 rows, unsorted, K up to N/2 (exact passes + ordered emit above the
    # LDS select)
    batch = 16
    qflags = rk.RK_LARGEST
    nq_max = 1 << 22
    q_ws_bytes = rk.workspace_size(rk.RK_F32, batch, batch * nq_max, nq_max // 2, qflags)
    qws = torch.empty(q_ws_bytes, dtype=torch.uint8, device="cuda")
    qov = torch.empty(batch * (nq_max // 2), dtype=torch.float32, device="cuda")
    qoi = torch.empty(batch * (nq_max // 2), dtype=torch.int32, device="cuda")
    t0 = time.perf_counter()
    quant = []
    for ln in range(16, 23):
        n = 1 << ln
        d_off = torch.arange(batch + 1, dtype=torch.int64, device="cuda") * n
        for k in (512, n // 100, n // 4, n // 2):
            d_k = torch.full((batch,), k, dtype=torch.int32, device="cuda")
            d_oo = torch.arange(batch + 1, dtype=torch.int64, device="cuda") * k

            def qcall():
                rc = L.rk_topk_batched(rk.RK_F32, vals.data_ptr(), d_off.data_ptr(),
                                       d_k.data_ptr(), batch, batch * n, k, qflags, None,
                                       qov.data_ptr(), qoi.data_ptr(), d_oo.data_ptr(),
                                       qws.data_ptr(), q_ws_bytes, sp)
                if rc != 0:
                    raise rk.RadikError(rc, "rk_topk_batched (quantile sweep)")

            ms = replay_ms(qcall)
            st = rk.status(qws, gstream)
            if st[0]:
                raise RuntimeError(f"quantile sweep N=2^{ln} K={k}: device status {st}")
            alg = batch * (n * 4 + k * 8 + 16)
            quant.append([n, k, round(ms, 5), batch * n / (ms / 1e3),
                          round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4), st[1], st[2]])
    log(f"sweep quantile: {len(quant)} points in {time.perf_counter() - t0:.1f} s")
    del vals, qws, qov, qoi
    return {
        "single_query": {
            "source": "radik eval/1-simple-topk.py:18-27 (batch 1)",
            "values": "U[0,1) fp32, seed 11, on the device", "largest": True, "sorted": True,
            "tunables": "defaults", "timing": "hipGraph replays, mean of 10",
            "columns": ["N", "K", "ms", "keys_per_s", "roofline_frac", "exact_path_ms",
                        "filter_gbs"],
            "points": single},
        "quantile": {
            "source": "radik eval/ex1-median.py:11-26 (batch 16, unsorted)",
            "values": "U[0,1) fp32, seed 11, on the device", "largest": True, "sorted": False,
            "tunables": "defaults", "timing": "hipGraph replays, mean of 10",
            "columns": ["N", "K", "ms", "keys_per_s", "roofline_frac", "exact_rows",
                        "ordered_rows"],
            "points": quant},
    }


def fallback_cost(args) -> dict:
    """Cost of the sampled path's fallback: 64 rows x 2^22 fp32, k = 1024,
    largest, sorted, default tunables, as N(0,1) rows (control) and as rows
    built against the sample (W.adversarial: "hidden" overflows the candidate
    slots, "decoy" leaves fewer than k keys above the threshold), so that
    every row takes the exact passes after its sampled attempt.  Also the
    exact path (RK_EXACT) on the control rows.  Each point: hipGraph replays
    (mean of 10), fallback rows from the device status, and the output
    checked against the oracle."""
    import numpy as np
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W
    from oracle import oracle as orc

    L = rk.lib()
    gstream = torch.cuda.Stream()
    sp = gstream.cuda_stream
    saved = rk.get_tune()
    rk.set_tune(tile=0, grid=0, sample_groups=0, group_elems=0, hist_bits=0, unroll=0)
    rows, n, k = 64, 1 << 22, 1024
    flags = rk.RK_LARGEST | rk.RK_SORTED
    out = {"rows": rows, "len": n, "k": k, "dtype": "f32", "tunables": "defaults",
           "timing": "hipGraph replays, mean of 10",
           "columns": ["ms", "keys_per_s", "roofline_frac", "fallback_rows", "bit_exact"]}
    for kind in ("control", "hidden", "decoy", "control_exact"):
        w = (W.adversarial(kind, rows=rows, n=n, k=k) if kind in ("hidden", "decoy") else
             W.Workload(kind, np.full(rows, n, np.int64), np.full(rows, k, np.int32),
                        True, 6, [W.Segment(0, rows, W.NORMAL, 0.0, 1.0)]))
        fl = flags | (rk.RK_EXACT if kind == "control_exact" else 0)
        vals = W.materialize(w)
        off = w.offsets()
        d_off = torch.from_numpy(off).cuda()
        d_k = torch.from_numpy(w.ks).cuda()
        d_oo = torch.from_numpy(w.out_offsets()).cuda()
        ws_bytes = rk.workspace_size(rk.RK_F32, rows, w.total, k, fl)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        ov = torch.empty(rows * k, dtype=torch.float32, device="cuda")
        oi = torch.empty(rows * k, dtype=torch.int32, device="cuda")

        def call():
            rc = L.rk_topk_batched(rk.RK_F32, vals.data_ptr(), d_off.data_ptr(), d_k.data_ptr(),
                                   rows, w.total, k, fl, None, ov.data_ptr(), oi.data_ptr(),
                                   d_oo.data_ptr(), ws.data_ptr(), ws_bytes, sp)
            if rc != 0:
                raise rk.RadikError(rc, "rk_topk_batched (fallback cost)")

        torch.cuda.synchronize()
        with rk.Graph.capture(gstream) as g:
            call()
        for _ in range(2):
            g.launch(gstream)
        gstream.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(gstream)
        for _ in range(10):
            g.launch(gstream)
        e1.record(gstream)
        gstream.synchronize()
        g.close()
        ms = e0.elapsed_time(e1) / 10
        st = rk.status(ws, gstream)
        bits = vals.view(torch.int32).cpu().numpy().view(np.uint32)
        eb, ei = orc.topk(bits, off, w.ks, flags=orc.LARGEST | orc.SORTED,
                          nthreads=orc.usable_threads())
        exact = bool(st[0] == 0 and np.array_equal(oi.cpu().numpy(), ei) and
                     np.array_equal(ov.view(torch.int32).cpu().numpy().view(np.uint32), eb))
        alg = w.algorithmic_bytes()
        out[kind] = [round(ms, 4), w.total / (ms / 1e3),
                     round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4), int(st[1]), exact]
        log(f"fallback cost {kind}: {ms:.3f} ms, fallback rows {st[1]}, bit-exact {exact}")
        del vals, ws, ov, oi, bits
        torch.cuda.empty_cache()
    rk.set_tune(**saved)
    out["hidden_vs_control"] = round(out["hidden"][0] / out["control"][0], 3)
    out["decoy_vs_control"] = round(out["decoy"][0] / out["control"][0], 3)
    return out


def knn_bench(args) -> dict:
    """Fused k-nearest-neighbour select (rk_knn_topk) on a resident bf16
    database against its unfused form: the [queries x database] fp32 scores
    written by rk_knn_scores, then rk_topk_batched over them.  Both give the
    same bits (checked here on the last call).  MFMA roofline: 2 * nq * n * d
    flops per call against the bf16 dense peak."""
    import torch

    import radik_amd as rk

    nq, n, d, k = (int(v) for v in args.knn_shape.split(","))
    g = torch.Generator(device="cuda").manual_seed(6)
    q = torch.randn(nq, d, device="cuda", generator=g).to(torch.bfloat16)
    x = torch.randn(n, d, device="cuda", generator=g).to(torch.bfloat16)
    flops = 2.0 * nq * n * d
    peak = 2500.0  # TFLOP/s, MI355X dense bf16
    steps, warm = max(args.steps, 1), max(args.warmup, 1)
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    out = {"workload": f"{nq} queries x {n} database vectors, d = {d}, bf16, k = {k}, sorted",
           "data": "synthetic: N(0,1) rounded to bf16 (torch generator seed 6)",
           "flops_per_call": flops, "database_bytes": n * d * 2, "steps": steps}
    scores = torch.empty((nq, n), dtype=torch.float32, device="cuda")
    off = torch.arange(nq + 1, dtype=torch.int64, device="cuda") * n
    ks = torch.full((nq,), k, dtype=torch.int32, device="cuda")
    for name, metric in (("inner_product", rk.RK_KNN_IP), ("l2", rk.RK_KNN_L2)):
        ws = torch.empty(rk.knn_workspace_size(rk.RK_BF16, nq, n, d, k), dtype=torch.uint8,
                         device="cuda")
        fv = torch.empty((nq, k), dtype=torch.float32, device="cuda")
        fi = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        largest = metric == rk.RK_KNN_IP
        flags = rk.RK_SORTED | (rk.RK_LARGEST if largest else 0)
        tws = torch.empty(rk.workspace_size(rk.RK_F32, nq, nq * n, k, flags), dtype=torch.uint8,
                          device="cuda")
        uv = torch.empty(nq * k, dtype=torch.float32, device="cuda")
        ui = torch.empty(nq * k, dtype=torch.int32, device="cuda")

        def fused():
            rk.knn_topk(q, x, k, metric=metric, sorted=True, workspace=ws, out_vals=fv,
                        out_idx=fi)

        def score_only():
            rk.knn_scores(q, x, metric=metric, out=scores)

        def unfused():
            score_only()
            rk.topk_batched(scores.view(-1), off, ks, largest=largest, sorted=True,
                            total_len=nq * n, max_k_=k, workspace=tws, out_vals=uv, out_idx=ui)

        f_ms = timed(fused)
        st = rk.status(ws)
        if st[0] != 0:
            raise RuntimeError(f"knn status {st}")
        s_ms = timed(score_only)
        u_ms = timed(unfused)
        same = bool(torch.equal(fv.view(-1).view(torch.int32), uv.view(torch.int32))
                    and torch.equal(fi.view(-1), ui))
        if not same:
            raise RuntimeError(f"knn {name}: fused and unfused results differ")
        out[name] = {
            "fused_ms": f_ms, "queries_per_s": nq / (f_ms / 1e3),
            "fused_tflops": flops / (f_ms / 1e3) / 1e12,
            "fused_frac_of_mfma_peak": flops / (f_ms / 1e3) / 1e12 / peak,
            "scores_only_ms": s_ms, "scores_tflops": flops / (s_ms / 1e3) / 1e12,
            "unfused_ms": u_ms, "fused_speedup": u_ms / f_ms,
            "second_pass_queries": st[1], "bit_exact_vs_unfused": same}
        log(f"knn {name}: fused {f_ms:.3f} ms ({out[name]['fused_tflops']:.0f} TFLOP/s), "
            f"scores only {s_ms:.3f} ms, unfused {u_ms:.3f} ms ({u_ms / f_ms:.2f}x)")
        del ws, tws
    out["roofline"] = {"bound": "mfma", "peak": peak, "unit": "TFLOP/s",
                       "achieved": out["inner_product"]["fused_tflops"],
                       "frac": out["inner_product"]["fused_frac_of_mfma_peak"]}
    return out


def cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo), for the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def percentile(x, q):
    """Nearest-rank percentile (the server's own statistic)."""
    import numpy as np
    x = np.sort(np.asarray(x))
    return float(x[max(1, int(np.ceil(q * len(x)))) - 1]) if len(x) else 0.0


def verify_requests(bodies, body, body_len, lens, ks, ov, oi, oo, every: int = 100) -> dict:
    """Checks every `every`-th request's output (values and indices, in the
    slots the replay wrote) against the oracle on the host; raises if any
    differs."""
    import numpy as np

    from oracle import oracle as orc
    sel = np.arange(0, len(lens), every)
    flat = bodies.view(-1).view(torch_int32())
    s_off = np.zeros(len(sel) + 1, np.int64)
    s_off[1:] = np.cumsum(lens[sel])
    host = np.empty(int(s_off[-1]), np.uint32)
    for j, i in enumerate(sel):
        b0 = int(body[i]) * body_len
        host[s_off[j]:s_off[j + 1]] = flat[b0:b0 + int(lens[i])].cpu().numpy().view(np.uint32)
    # each request has k slots (results then the tail fill), min(len, k) results
    eb, ei = orc.topk(host, s_off, ks[sel], flags=orc.LARGEST | orc.SORTED,
                      nthreads=orc.usable_threads())
    m = np.minimum(lens[sel], ks[sel].astype(np.int64))
    e_off = np.zeros(len(sel) + 1, np.int64)
    e_off[1:] = np.cumsum(m)
    gv = ov.view(torch_int32()).cpu().numpy().view(np.uint32)
    gi = oi.cpu().numpy()
    for j, i in enumerate(sel):
        a = int(oo[i])
        mm = int(m[j])
        if not (np.array_equal(gv[a:a + mm], eb[e_off[j]:e_off[j + 1]])
                and np.array_equal(gi[a:a + mm], ei[e_off[j]:e_off[j + 1]])):
            raise RuntimeError(f"serving: request {i} differs from the oracle")
    return {"requests": int(len(sel)), "every": every, "bit_exact": True,
            "against": "oracle (CPU restatement), values and indices"}


def torch_int32():
    import torch
    return torch.int32


def serving(args) -> dict:
    """Serving-loop measurement.  Request stream: n requests with length
    floor(2^U(12,22)), k floor(2^U(0,12)) and one of four value families
    (uniform, normal, Zipf s=1.1, near-constant with 30% duplicates); each
    request reads a prefix of a resident 2^22-element body of its family
    (256 bodies per family, 16 GB).  Poisson arrivals at `load` times the
    plain server's measured capacity; the continuous-batching server and the
    plain server (one RK_EXACT call per request, one stream) replay the same
    schedule against the same 32 GB admission budget."""
    import numpy as np
    import torch

    import radik_amd as rk
    import radik_amd.workloads as W

    n = args.serve_requests
    lens, ks, kinds = W.c5_requests(n)
    per_kind, body_len = 256, 1 << 22
    bodies = W.c5_bodies(per_kind=per_kind, body_len=body_len)
    body = kinds.astype(np.int64) * per_kind + (np.arange(n) % per_kind)
    vptr = bodies.data_ptr() + body * body_len * 4
    oo = np.zeros(n + 1, np.int64)
    oo[1:] = np.cumsum(ks.astype(np.int64))
    ov = torch.empty(int(oo[-1]), dtype=torch.float32, device="cuda")
    oi = torch.empty(int(oo[-1]), dtype=torch.int32, device="cuda")
    optr_v = ov.data_ptr() + oo[:-1] * 4
    optr_i = oi.data_ptr() + oo[:-1] * 4
    torch.cuda.synchronize()
    log(f"serving: {n} requests, mean len {lens.mean():.0f}, mean k {ks.mean():.0f}, "
        f"bodies {bodies.numel() * 4 / 1e9:.1f} GB resident")
    budget = 32 << 30
    common = dict(max_k=4096, mem_budget=budget, max_batch_keys=1 << 27)
    batched_cfg = dict(max_batch_rows=1024, streams=4, **common)

    def run(plain, idx, arrivals):
        cfg = dict(plain=True, **common) if plain else batched_cfg
        with rk.Server(**cfg) as srv:
            lat, makespan = srv.replay(vptr[idx], lens[idx], ks[idx], optr_v[idx], optr_i[idx],
                                       arrivals)
            st = srv.stats()
        if st["failed"] or st["rejected"] or np.any(lat < 0):
            raise RuntimeError(f"serving run failed: {st}")
        return lat, makespan, st

    # warm both paths (module load, LDS attributes, resident-grid queries)
    warm = np.arange(min(n, 256))
    run(True, warm, np.zeros(len(warm)))
    run(False, warm, np.zeros(len(warm)))
    # capacity: every request of a prefix submitted at t = 0 (closed burst)
    n_cap = min(n, 4000)
    cap_idx = np.arange(n_cap)
    _, mk_plain, _ = run(True, cap_idx, np.zeros(n_cap))
    cap_plain = n_cap / mk_plain
    n_cap_b = min(n, 40000)
    _, mk_b, st_cap = run(False, np.arange(n_cap_b), np.zeros(n_cap_b))
    cap_batched = n_cap_b / mk_b
    log(f"capacity: plain {cap_plain:.0f} req/s, batched {cap_batched:.0f} req/s "
        f"(mean batch {st_cap['mean_batch_rows']:.1f} rows)")
    rate = args.serve_load * cap_plain
    arrivals = W.poisson_arrivals(n, rate)
    idx = np.arange(n)
    out = {"requests": n, "offered_rate_per_s": rate, "load_vs_plain_capacity": args.serve_load,
           "mem_budget_bytes": budget,
           "workload": "len floor(2^U(12,22)), k floor(2^U(0,12)), fp32 mixture of uniform / "
                       "normal / Zipf(1.1) / near-constant+30% duplicates, Poisson arrivals",
           "latency": "scheduled arrival to observed completion, us"}
    for name, plain in (("batched", False), ("plain", True)):
        lat, makespan, st = run(plain, idx, arrivals)
        out[name] = {"p50_us": percentile(lat, 0.50), "p99_us": percentile(lat, 0.99),
                     "mean_us": float(lat.mean()), "max_us": float(lat.max()),
                     "throughput_req_s": n / makespan,
                     "keys_per_s": float(lens.sum()) / makespan,
                     "batches": st["batches"], "mean_batch_rows": st["mean_batch_rows"],
                     "fallback_batches": st["fallback_batches"],
                     "peak_inflight_bytes": st["peak_inflight_bytes"],
                     "server_bytes": st["server_bytes"]}
        log(f"{name}: p50 {out[name]['p50_us']:.0f} us, p99 {out[name]['p99_us']:.0f} us, "
            f"{out[name]['throughput_req_s']:.0f} req/s, mean batch {st['mean_batch_rows']:.2f}")
    out["batched"]["capacity_req_s"] = cap_batched
    out["plain"]["capacity_req_s"] = cap_plain
    out["p99_ratio_plain_over_batched"] = out["plain"]["p99_us"] / out["batched"]["p99_us"]

    # Every 100th request of the last replay (the plain server's, which wrote
    # the same slots) against the oracle on the host.
    out["plain"]["verified"] = verify_requests(bodies, body, body_len, lens, ks, ov, oi, oo)

    # A load where batches form: 0.5 x the batched server's own capacity (the
    # plain server cannot follow at that rate, so it is not replayed there).
    hi_rate = args.serve_hi_load * cap_batched
    hi_arrivals = W.poisson_arrivals(n, hi_rate, seed=6)
    lat, makespan, st = run(False, idx, hi_arrivals)
    out["batched_high_load"] = {
        "offered_rate_per_s": hi_rate, "load_vs_batched_capacity": args.serve_hi_load,
        "p50_us": percentile(lat, 0.50), "p99_us": percentile(lat, 0.99),
        "mean_us": float(lat.mean()), "max_us": float(lat.max()),
        "throughput_req_s": n / makespan, "keys_per_s": float(lens.sum()) / makespan,
        "batches": st["batches"], "mean_batch_rows": st["mean_batch_rows"],
        "max_batch_rows": st["max_batch_rows"], "fallback_batches": st["fallback_batches"],
        "peak_inflight_bytes": st["peak_inflight_bytes"],
        "verified": verify_requests(bodies, body, body_len, lens, ks, ov, oi, oo)}
    log(f"batched at {hi_rate:.0f} req/s: p50 {out['batched_high_load']['p50_us']:.0f} us, "
        f"p99 {out['batched_high_load']['p99_us']:.0f} us, mean batch "
        f"{st['mean_batch_rows']:.1f}, verified {out['batched_high_load']['verified']}")
    del bodies, ov, oi
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    sys.exit(main())
