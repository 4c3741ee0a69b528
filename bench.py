"""Benchmark of the windowed stream joins on one MI355X.

Headline (the JSON line's metric, value, roofline, cpu_baseline): lazy HIP
hash join over tumbling count windows of 2^22 tuples per side, two streams of
2^26 16-byte tuples (keys uniform in [0, 2^24), seeds 123/456, val
independent uniform, ts = index): 16 windows per step, count + C1 + C2 per
window, inputs resident in HBM before the timed region starts.
  value        joined input tuples/s = 2 * 2^26 * steps / time
  roofline     HBM roofline of the dominant phase (algorithmic bytes per
               launch / average launch time from fence-free HIP events on
               the ctx stream); traffic is null (HBM bytes come from
               rocprofv3 --pmc runs of this build, not from this process)
  window_ms    p50 / p99 device time per window
  cpu_baseline the C restatement of the reference CPU hash join (oracle)
               on the same 16 windows, INTRA (1 joiner x P probe threads)
               and INTER (P joiners), median of 5, P = the CPUs this process
               may run on

The default run also measures every other workload of the scope under the
same clock and puts them in "extra" (each with ms_per_step, a roofline, an
oracle check of sampled results and a CPU baseline):
  cfg1     CPU HJ at 2^16-tuple windows over 2^20-tuple streams, keys
           U[0, 2^16) (INTRA and INTER, median of 5), the HIP HJ on the
           same windows, and window 0's full pair list against the oracle
  nlj      lazy NLJ at 2^16-tuple windows, keys U[0, 2^24)
  batch    4096 lognormal windows per launch (median 2^14, sigma 1, clipped
           [2^8, 2^20]), Zipf theta in {0, 0.8, 1.2} over [0, 2^28), HJ and
           SMJ, autotuned and replayed as a captured plan
  learned  learned vs radix partitioning, 8 windows of 2^22 per side,
           drifting GMM + 30% Zipf 1.2 keys, same drift epoch and right
           after a drift step
  slide    eager sliding window W = 2^22, events of Poisson(2^16) tuples,
           Zipf 1.0 over [0, 2^26): throughput and p50/p99 per-event latency

--workload X runs one of them alone (window = the headline only) and prints
its own JSON line; eager is only reachable that way.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-stream-join-benchmark_amd", "py"))

HBM_PEAK = 8.0e12
# window path: one window in TIME_EVERY carries the three timestamp events
# the per-phase kernel times and latency percentiles come from (40 timed
# windows over 10 steps); each timed window costs the pipeline ~6 us of
# event overhead, so sampling fewer keeps the measured throughput close to
# the uninstrumented one
TIME_EVERY = 4
VALU_PEAK_OPS = 256 * 64 * 2.4e9  # MI355X: 256 CUs x 64 int32 lanes per clock x 2.4 GHz
NLJ_OPS_PER_COMPARE = 1.0         # the algorithm's own count: one key compare per pair


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--alg", default="HJ", choices=["HJ", "SMJ"])
    ap.add_argument("--log-window", type=int, default=22)
    ap.add_argument("--log-stream", type=int, default=26)
    ap.add_argument("--key-range-log", type=int, default=24)
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--overlap", action="store_true",
                    help="window: run window w+1's partition pass on a second stream, concurrently with "
                         "window w's join (per-kernel durations then overlap)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baselines: median of this many runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 128 MiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--workload", default="all",
                    choices=["all", "window", "cfg1", "batch", "learned", "slide", "nlj", "eager"],
                    help="all (default): the headline window line with every other workload in 'extra'")
    ap.add_argument("--theta", type=float, default=None, help="batch: only this Zipf theta (default 0, 0.8, 1.2)")
    ap.add_argument("--n-windows", type=int, default=4096, help="batch: windows per launch")
    ap.add_argument("--no-autotune", action="store_true", help="batch: keep the default launch shape")
    ap.add_argument("--extra-steps", type=int, default=10, help="timed steps of each extra workload")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baselines (the oracle: test infrastructure, used here only as the timed
# CPU reference)
# ---------------------------------------------------------------------------
def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_info():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    quota = None
    try:
        q = open("/sys/fs/cgroup/cpu.max").read().split()
        if q[0] != "max":
            quota = float(q[0]) / float(q[1])
    except Exception:
        pass
    return {"model": model, "hardware_concurrency": os.cpu_count(), "affinity_cpus": cpu_threads(),
            "cgroup_cpu_quota": quota}


def oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.lib()
    return oracle


def median_of(fn, runs):
    import numpy as np
    secs, last = [], None
    for _ in range(max(runs, 1)):
        s, last = fn()
        secs.append(s)
    return float(np.median(secs)), secs, last


def cpu_lazy_hj(R, S, W, D, runs, label):
    """The restated CPU_HJ over R/S's tumbling windows of W tuples, INTRA and
    INTER with P threads, median of `runs`."""
    orc = oracle_mod()
    P = cpu_threads()
    n = min(len(R), len(S)) // W * W
    out = {"cores": P, "cpu": cpu_info(), "windows": n // W, "window": W, "buckets_D": D, "runs": runs}
    for name, inter in (("INTRA", False), ("INTER", True)):
        med, secs, tot = median_of(lambda: orc.bench_lazy_hj(R[:n], S[:n], W, D, P, inter), runs)
        out[name] = {"tuples_per_s": 2 * n / med, "median_s": med, "runs_s": secs,
                     "count": int(tot.count), "c1": int(tot.c1), "c2": int(tot.c2)}
        log(f"cpu {label} {name} P={P}: {2 * n / med / 1e6:.1f} M tuples/s")
    best = max(("INTRA", "INTER"), key=lambda k: out[k]["tuples_per_s"])
    out["best"] = best
    return out


def cpu_windows_pool(fn, jobs):
    """Runs fn(job) for every job on P threads (the oracle's ctypes calls
    release the GIL: INTER, one window per thread at a time)."""
    from concurrent.futures import ThreadPoolExecutor
    P = cpu_threads()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=P) as ex:
        res = list(ex.map(fn, jobs))
    return time.perf_counter() - t0, res


# ---------------------------------------------------------------------------
# headline: lazy HJ (or SMJ) windows of 2^22 tuples per side
# ---------------------------------------------------------------------------
def run_window(a, np, torch, sjb):
    W = 1 << a.log_window
    N = 1 << a.log_stream
    nw = N // W
    K = 1 << a.key_range_log
    alg = sjb.ALG[a.alg]
    flags = (0 if a.no_c2 else sjb.F_C2) | (sjb.F_OVERLAP if a.overlap else 0)
    # every TIME_EVERY-th window carries the library's timestamp events
    # (three fence-free events per window): per-phase kernel times and the
    # window latency percentiles come from those windows, at a fraction of
    # the events' cost to the measured throughput
    n_timed = (nw + TIME_EVERY - 1) // TIME_EVERY

    t0 = time.time()
    R = sjb.generate(N, sjb.DIST_UNIFORM, K, 0.0, seed=123)
    S = sjb.generate(N, sjb.DIST_UNIFORM, K, 0.0, seed=456)
    dR = torch.from_numpy(R.view(np.uint8)).to("cuda")
    dS = torch.from_numpy(S.view(np.uint8)).to("cuda")
    log(f"window: generated + uploaded 2 x {N} tuples in {time.time() - t0:.1f}s")

    ctx = sjb.Context(device=0, max_window=W, distinct_keys=K)
    ctx.timing_ring(n_timed * a.steps)
    stream = torch.cuda.ExternalStream(ctx.stream)
    d_res = torch.zeros(nw * 32, dtype=torch.uint8, device="cuda")
    base_r, base_s, base_res = dR.data_ptr(), dS.data_ptr(), d_res.data_ptr()

    def step(evs=None):
        if evs is not None:
            evs[0].record(stream)
        for w in range(nw):
            f = flags | (sjb.F_TIMING if w % TIME_EVERY == 0 else 0)
            ctx.lazy_join_async(alg, f, base_r + w * W * 16, W, base_s + w * W * 16, W, base_res + w * 32)
        if evs is not None:
            evs[1].record(stream)

    for _ in range(a.warmup):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    first = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()

    # step boundaries only; per-window and per-phase times come from the
    # library's timestamp ring
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(a.steps)]
    torch.cuda.synchronize()
    ctx.sync()
    t_start = time.perf_counter()
    for s in range(a.steps):
        step(evs[s])
    ctx.sync()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_start

    dev_s = sum(evs[s][0].elapsed_time(evs[s][1]) for s in range(a.steps)) / 1e3
    part_ms, join_ms = ctx.timing_read(n_timed * a.steps)
    win_ms = part_ms + join_ms
    last = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()
    if a.warmup > 0 and not np.array_equal(first, last):
        raise SystemExit("window results changed between steps")
    ctx.close()

    tuples = 2 * N * a.steps
    # Algorithmic bytes per launch: the partition pass reads both sides'
    # tuples (16 B) and writes one 8-byte record per tuple; the join phase
    # (per-partition join + reduction) reads every record once.
    phases = {"partition": {"alg_bytes": 24 * 2 * W, "ms": float(part_ms.mean()),
                            "kernels": "partition pass (k_partition_pipe; SMJ: + key-range sample)"},
              "join": {"alg_bytes": 8 * 2 * W, "ms": float(join_ms.mean()),
                       "kernels": "k_join_pipe (resident grid; sums into the window result, no reduction "
                                  "kernel)"}}
    dom = max(phases, key=lambda k: phases[k]["ms"])
    for v in phases.values():
        v["achieved_GBps"] = v["alg_bytes"] / (v["ms"] / 1e3) / 1e9
        v["frac"] = v["achieved_GBps"] * 1e9 / HBM_PEAK
    pipe_bytes = 16 * 2 * W
    avg_win = float(win_ms.mean()) / 1e3
    out = {
        "metric": "joined input tuples/s + %HBM roofline at 2^22-tuple windows; p99 window latency",
        "value": tuples / t_wall, "unit": "tuples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * t_wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 keys / u64 sums",
        "data": "synthetic (splitmix64 generator, uniform keys)",
        "config": {"workload": f"lazy {a.alg} tumbling windows 2^{a.log_window}/side, streams "
                               f"2^{a.log_stream} tuples, keys U[0,2^{a.key_range_log})",
                   "windows_per_step": nw, "checksums": "count+C1" + ("" if a.no_c2 else "+C2"),
                   "parallelism": "single GPU",
                   "overlap": "partition of window w+1 concurrent with join of window w (2 streams)"
                   if a.overlap else "none"},
        "device_time_s": dev_s,
        "window_ms": {"p50": float(np.percentile(win_ms, 50)), "p99": float(np.percentile(win_ms, 99)),
                      "mean": float(win_ms.mean()), "timed_windows": int(len(win_ms)),
                      "sampling": f"1 in {TIME_EVERY} windows carries timestamp events"},
        "roofline": {"bound": "hbm", "kernel": dom, "kernels_timed": phases[dom]["kernels"],
                     "achieved": phases[dom]["achieved_GBps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": phases[dom]["frac"], "traffic": None,
                     "alg_bytes_per_launch": phases[dom]["alg_bytes"], "avg_launch_ms": phases[dom]["ms"]},
        "phases": phases,
        "pipeline_roofline": {"alg_bytes_per_window": pipe_bytes, "achieved": pipe_bytes / avg_win / 1e9,
                              "frac": pipe_bytes / avg_win / HBM_PEAK},
        "result_window0": {"count": int(last[0]["count"]), "c1": hex(int(last[0]["checksum1"]))},
    }
    # achievable streaming rate for reference: a plain 1 GiB device copy
    # (read + write), timed with events after the measured steps
    cbuf = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
    cdst = torch.empty_like(cbuf)
    cdst.copy_(cbuf)
    ce = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ce[0].record()
    for _ in range(5):
        cdst.copy_(cbuf)
    ce[1].record()
    torch.cuda.synchronize()
    copy_bps = 5 * 2 * (1 << 30) / (ce[0].elapsed_time(ce[1]) / 1e3)
    del cbuf, cdst
    for v in phases.values():
        v["frac_of_copy"] = v["achieved_GBps"] * 1e9 / copy_bps
    out["copy_GBps"] = copy_bps / 1e9
    if a.calibrate_copy:
        src = torch.empty(128 << 20, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = 128 << 20
    del dR, dS
    torch.cuda.empty_cache()
    if not a.no_cpu_baseline:
        cb = cpu_lazy_hj(R, S, W, 1 << 20, a.cpu_runs, "cfg2")
        best = cb[cb["best"]]
        gpu_count = int(last["count"].sum())
        gpu_c1 = int(last["checksum1"].astype(np.uint64).sum(dtype=np.uint64))
        out["cpu_baseline"] = {
            "value": best["tuples_per_s"], "unit": "tuples/s", "cores": cb["cores"], "kind": "port",
            "sample": f"all {nw} windows of the step (2^{a.log_window}/side), lazy CPU_HJ restated in C, "
                      f"D=2^20 buckets, {cb['best']} decomposition (faster of INTRA 1 joiner x {cb['cores']} "
                      f"probe threads and INTER {cb['cores']} joiners), median of {a.cpu_runs} runs",
            "cpu": cb["cpu"], "INTRA_tuples_per_s": cb["INTRA"]["tuples_per_s"],
            "INTER_tuples_per_s": cb["INTER"]["tuples_per_s"],
            "count_matches_gpu": best["count"] == gpu_count, "c1_matches_gpu": best["c1"] == gpu_c1}
    return out, R, S


# ---------------------------------------------------------------------------
# cfg1: the reference's CPU-runnable configuration
# ---------------------------------------------------------------------------
def run_cfg1(a, np, torch, sjb):
    W, N = 1 << 16, 1 << 20
    nw = N // W
    R = sjb.generate(N, sjb.DIST_UNIFORM, 1 << 16, 0.0, seed=123)
    S = sjb.generate(N, sjb.DIST_UNIFORM, 1 << 16, 0.0, seed=456)
    orc = oracle_mod()
    cpu = cpu_lazy_hj(R, S, W, 1 << 16, a.cpu_runs, "cfg1")
    dR = torch.from_numpy(R.view(np.uint8)).to("cuda")
    dS = torch.from_numpy(S.view(np.uint8)).to("cuda")
    ctx = sjb.Context(device=0, max_window=W, distinct_keys=1 << 16)
    stream = torch.cuda.ExternalStream(ctx.stream)
    d_res = torch.zeros(nw * 32, dtype=torch.uint8, device="cuda")

    def step():
        for w in range(nw):
            ctx.lazy_join_async(sjb.ALG["HJ"], sjb.F_C2, dR.data_ptr() + w * W * 16, W,
                                dS.data_ptr() + w * W * 16, W, d_res.data_ptr() + w * 32)

    step()
    ctx.sync()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    steps = 5 * a.extra_steps
    ev[0].record(stream)
    for _ in range(steps):
        step()
    ev[1].record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / steps
    res = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()
    # per-window oracle check and window 0's full pair list
    per_window_ok = True
    for w in range(nw):
        o, _ = orc.lazy_hj(R[w * W:(w + 1) * W], S[w * W:(w + 1) * W], 1 << 16)
        per_window_ok &= (int(res[w]["count"]), int(res[w]["checksum1"]), int(res[w]["checksum2"])) == o.triple()
    o0, pairs = orc.lazy_hj(R[:W], S[:W], 1 << 16, want_pairs=True, cap=4 * W)
    cap = int(o0.count) + 16
    out_buf = torch.zeros(cap * 32, dtype=torch.uint8, device="cuda")
    st = ctx.lazy_join(sjb.ALG["HJ"], sjb.OUT["CountKernel"], 0, dR, W, dS, W, out_buf, cap)
    rows = out_buf.cpu().numpy().view(orc.PAIR_DTYPE)[: st.pairs_written]
    pairs_ok = bool(np.array_equal(orc.sort_pairs(rows), orc.sort_pairs(pairs)))
    ctx.close()
    tot_count = int(res["count"].sum())
    return {"workload": "lazy HJ tumbling windows 2^16/side over 2^20-tuple streams, keys U[0,2^16), "
                        "count+C1+C2 (+ window 0's full pair list)",
            "gpu": {"ms_per_step": ms, "tuples_per_s": 2 * N / (ms / 1e3),
                    "roofline": {"bound": "hbm", "kernel": "whole window (partition + join + reduce)",
                                 "achieved": 16 * 2 * N / (ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9,
                                 "unit": "GB/s", "frac": 16 * 2 * N / (ms / 1e3) / HBM_PEAK}},
            "cpu_baseline": {"value": cpu[cpu["best"]]["tuples_per_s"], "unit": "tuples/s", "cores": cpu["cores"],
                             "kind": "port", "sample": f"the whole configuration ({nw} windows), median of "
                                                       f"{a.cpu_runs} runs, {cpu['best']}",
                             "INTRA_tuples_per_s": cpu["INTRA"]["tuples_per_s"],
                             "INTER_tuples_per_s": cpu["INTER"]["tuples_per_s"], "cpu": cpu["cpu"]},
            "oracle_check": {"windows_equal": bool(per_window_ok), "window0_pairs": int(st.pairs_written),
                             "window0_pairs_equal": pairs_ok,
                             "cpu_totals_equal": cpu["INTER"]["count"] == tot_count},
            "total_matches": tot_count}


# ---------------------------------------------------------------------------
# nested-loop join at 2^16-tuple windows
# ---------------------------------------------------------------------------
def nlj_bitmap_windows(np, torch, sjb, ctx, R, S, br, bs, W, nw, counts):
    """The Bitmap method (pair rows through per-left-tuple 64-key match
    words) on the first nw windows: one synchronous sjb_lazy_join per window
    (mask pass, mask count read back, device decode), timed by the host clock
    around each call; every window's count must equal the count kernel's,
    and window 0's rows the oracle's."""
    cap = int(counts["count"][:nw].max()) + 64
    out = torch.zeros(cap * 32, dtype=torch.uint8, device="cuda")
    nlj, bm = sjb.ALG["NLJ"], sjb.OUT["Bitmap"]
    ctx.lazy_join(nlj, bm, 0, br, W, bs, W, out, cap)  # warm-up
    ms, same = [], True
    for w in range(nw):
        t0 = time.perf_counter()
        st = ctx.lazy_join(nlj, bm, 0, br + w * W * 16, W, bs + w * W * 16, W, out, cap)
        ms.append(1e3 * (time.perf_counter() - t0))
        same &= st.result_count == int(counts["count"][w]) and st.pairs_written == st.result_count
        if w == 0:
            rows = out.cpu().numpy()[: 32 * st.pairs_written].copy()
    orc = oracle_mod()
    o, pairs = orc.lazy_hj(R[:W], S[:W], 1 << 16, want_pairs=True, cap=cap)
    rows = rows.view(orc.PAIR_DTYPE)
    ms = np.array(ms)
    return {"windows": nw, "method": "Bitmap (masks + device decode into ResultTuple rows)",
            "ms_per_window_p50": float(np.percentile(ms, 50)), "compares_per_s": float(W) * W / (float(ms.mean()) / 1e3),
            "note": "host clock around each synchronous call (two host syncs per window)",
            "counts_equal_count_kernel": bool(same),
            "window0_rows_equal_oracle": bool(np.array_equal(orc.sort_pairs(rows), orc.sort_pairs(pairs)))}


def run_nlj(a, np, torch, sjb, steps, warmup):
    W, nw = 1 << 16, 64
    N = W * nw
    R = sjb.generate(N, sjb.DIST_UNIFORM, 1 << 24, 0.0, seed=123)
    S = sjb.generate(N, sjb.DIST_UNIFORM, 1 << 24, 0.0, seed=456)
    dR = torch.from_numpy(R.view(np.uint8)).to("cuda")
    dS = torch.from_numpy(S.view(np.uint8)).to("cuda")
    ctx = sjb.Context(device=0, max_window=W, distinct_keys=1 << 24)
    stream = torch.cuda.ExternalStream(ctx.stream)
    d_res = torch.zeros(nw * 32, dtype=torch.uint8, device="cuda")
    br, bs, bo = dR.data_ptr(), dS.data_ptr(), d_res.data_ptr()
    flags = 0 if a.no_c2 else sjb.F_C2

    def step(evs=None):
        for w in range(nw):
            if evs is not None:
                evs[w][0].record(stream)
            ctx.lazy_join_async(sjb.ALG["NLJ"], flags, br + w * W * 16, W, bs + w * W * 16, W, bo + w * 32)
            if evs is not None:
                evs[w][1].record(stream)

    for _ in range(warmup):
        step()
    ctx.sync()
    first = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()
    # every window of the last step carries events (window = key compaction
    # + compare); the other steps run without them
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(nw)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s_ in range(steps):
        step(evs if s_ == steps - 1 else None)
    ctx.sync()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_start
    win_ms = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    last = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()
    bitmap = nlj_bitmap_windows(np, torch, sjb, ctx, R, S, br, bs, W, min(nw, 16), last)
    ctx.close()
    if warmup > 0 and not np.array_equal(first, last):
        raise SystemExit("NLJ results changed between steps")
    compares = float(W) * W
    avg = float(win_ms.mean()) / 1e3
    achieved = compares * NLJ_OPS_PER_COMPARE / avg
    orc = oracle_mod()
    check = []
    for w in (0, nw - 1):
        o, _ = orc.lazy_hj(R[w * W:(w + 1) * W], S[w * W:(w + 1) * W], 1 << 16)
        check.append((int(last[w]["count"]), int(last[w]["checksum1"]), int(last[w]["checksum2"])) == o.triple())
    out = {
        "metric": "NLJ compares/s (lazy nested-loop join, 2^16-tuple windows)",
        "value": compares * nw * steps / t_wall, "unit": "compares/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * t_wall / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 keys / u64 sums", "data": "synthetic (splitmix64 generator, uniform keys)",
        "config": {"workload": f"lazy NLJ tumbling windows 2^16/side, {nw} windows per step, keys U[0,2^24)",
                   "checksums": "count+C1" + ("" if a.no_c2 else "+C2"), "parallelism": "single GPU"},
        "input_tuples_per_s": 2 * N * steps / t_wall,
        "window_ms": {"p50": float(np.percentile(win_ms, 50)), "p99": float(np.percentile(win_ms, 99)),
                      "mean": float(win_ms.mean())},
        "roofline": {"bound": "valu", "kernel": "NLJ window (key compaction + compare; workgroups sum into the result)",
                     "achieved": achieved / 1e12, "peak": VALU_PEAK_OPS / 1e12, "unit": "T compares/s",
                     "frac": achieved / VALU_PEAK_OPS, "traffic": None,
                     "note": "one int32 compare (one v_cmpx_ne_u32 lane) per pair against one VALU lane-op per clock",
                     "ops_per_launch": compares * NLJ_OPS_PER_COMPARE, "avg_launch_ms": 1e3 * avg},
        "oracle_check": {"windows_checked": [0, nw - 1], "equal": all(check)},
        "total_matches": int(last["count"].sum()),
        "bitmap": bitmap,
    }
    if not a.no_cpu_baseline:
        P = cpu_threads()
        jobs = [(R[w * W:(w + 1) * W], S[w * W:(w + 1) * W]) for w in range(min(nw, 2 * P))]
        secs, res = cpu_windows_pool(lambda j: orc.lazy_nlj(j[0], j[1])[0], jobs)
        ok = all((int(r.count), int(r.c1)) == (int(last[w]["count"]), int(last[w]["checksum1"]))
                 for w, r in enumerate(res))
        out["cpu_baseline"] = {"value": compares * len(jobs) / secs, "unit": "compares/s", "cores": P,
                               "kind": "port", "sample": f"windows 0..{len(jobs) - 1} (2^16 x 2^16 each), lazy "
                                                         f"CPU_NLJ restated in C, one window per thread "
                                                         f"(INTER, {P} threads)",
                               "seconds": secs, "results_match_gpu": ok, "cpu": cpu_info()}
    return out


# ---------------------------------------------------------------------------
# batched uneven windows (one persistent launch per step, captured plan)
# ---------------------------------------------------------------------------
def lognormal_sizes(n, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    sz = np.exp(rng.normal(np.log(2.0 ** 14), 1.0, size=n))
    return np.clip(np.round(sz), 2 ** 8, 2 ** 20).astype(np.uint64)


def time_plan(ctx, plan, steps, warmup, torch):
    stream = torch.cuda.ExternalStream(ctx.stream)
    for _ in range(warmup):
        plan.launch()
    ctx.sync()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for s in range(steps):
        plan.launch()
        ev[s + 1].record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    return [ev[s].elapsed_time(ev[s + 1]) for s in range(steps)], time.perf_counter() - t0


def run_batch_theta(a, np, torch, sjb, theta, algs, steps, warmup):
    """4096 lognormal windows at one Zipf theta: HJ and SMJ on the same data."""
    nwin = a.n_windows
    nl, nr = lognormal_sizes(nwin, 11), lognormal_sizes(nwin, 12)
    lo = np.concatenate([[0], np.cumsum(nl)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(nr)]).astype(np.uint64)
    dist = sjb.DIST_ZIPF if theta > 0 else sjb.DIST_UNIFORM
    t0 = time.time()
    R = sjb.generate(int(lo[-1]), dist, 1 << 28, theta, seed=123)
    S = sjb.generate(int(ro[-1]), dist, 1 << 28, theta, seed=456)
    dR = torch.from_numpy(R.view(np.uint8)).to("cuda")
    dS = torch.from_numpy(S.view(np.uint8)).to("cuda")
    log(f"batch theta={theta}: {nwin} windows, {int(lo[-1])} + {int(ro[-1])} tuples in {time.time() - t0:.1f}s")
    grouped = theta >= 1.0
    flags = sjb.F_GROUPED if grouped else (0 if a.no_c2 else sjb.F_C2)
    tuples = int(lo[-1] + ro[-1])
    alg_bytes = 16 * tuples
    orc = oracle_mod()
    rng = np.random.default_rng(int(theta * 10) + 3)
    order = np.argsort(nl + nr)
    pick = sorted(set(order[-4:].tolist()) | set(rng.choice(nwin, 4, replace=False).tolist()))
    ref = {}
    for w in pick:
        r, s = R[lo[w]:lo[w + 1]], S[ro[w]:ro[w + 1]]
        if grouped:
            o, _ = orc.lazy_smj(r, s)
            ref[w] = (o.count, o.c1)
        else:
            o, _ = orc.lazy_hj(r, s, 1 << max(10, int(len(r)).bit_length()))
            ref[w] = o.triple()
    out = {}
    ctx = sjb.Context(device=0, max_window=1 << 20, distinct_keys=1 << 20)
    for name in algs:
        alg = sjb.ALG[name]
        tune = None
        if not a.no_autotune:
            t1 = time.time()
            best, tried = ctx.batch_autotune(alg, flags, dR, lo, dS, ro, reps=2)
            tune = dict(best.as_dict(), tried=tried, seconds=time.time() - t1)
        d_res = torch.zeros(nwin * 32, dtype=torch.uint8, device="cuda")
        plan = ctx.batch_plan(alg, flags, dR, lo, dS, ro, d_res)
        info = plan.info()
        ms, wall = time_plan(ctx, plan, steps, warmup, torch)
        res = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()
        plan.destroy()
        ok = True
        for w in pick:
            x = res[w]
            got = (int(x["count"]), int(x["checksum1"])) if grouped else \
                (int(x["count"]), int(x["checksum1"]), int(x["checksum2"]))
            ok &= got == ref[w]
        avg = float(np.mean(ms)) / 1e3
        out[name] = {"ms_per_step": 1e3 * wall / steps, "tuples_per_s": tuples * steps / wall,
                     "step_ms": {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99))},
                     "roofline": {"bound": "hbm", "kernel": "batch pipeline (clear + partition + persistent join)",
                                  "achieved": alg_bytes / avg / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                  "frac": alg_bytes / avg / HBM_PEAK, "traffic": None,
                                  "alg_bytes_per_launch": alg_bytes},
                     "plan": {"n_chunks": info.n_chunks, "n_items": info.n_items, "blocks": info.blocks},
                     "autotune": tune, "oracle_check": {"windows": [int(w) for w in pick], "equal": bool(ok)},
                     "total_matches": int(res["count"].sum())}
        log(f"batch theta={theta} {name}: {out[name]['tuples_per_s'] / 1e9:.1f} G tuples/s, oracle ok={ok}")
    ctx.close()
    del dR, dS
    torch.cuda.empty_cache()
    if not a.no_cpu_baseline:
        # every window of the launch on P threads, one window per thread at a
        # time: CPU_HJ with C2 (theta < 1), the closed-form CPU_SMJ count +
        # C1 at theta 1.2 (pairs are not enumerable there)
        if grouped:
            fn = lambda w: orc.lazy_smj(R[lo[w]:lo[w + 1]], S[ro[w]:ro[w + 1]])[0]
        else:
            fn = lambda w: orc.lazy_hj(R[lo[w]:lo[w + 1]], S[ro[w]:ro[w + 1]],
                                       1 << max(10, int(lo[w + 1] - lo[w]).bit_length()))[0]
        secs, res_cpu = cpu_windows_pool(fn, range(nwin))
        cpu_count = sum(int(r.count) for r in res_cpu)
        out["cpu_baseline"] = {"value": tuples / secs, "unit": "tuples/s", "cores": cpu_threads(), "kind": "port",
                               "sample": f"all {nwin} windows, " + ("CPU_SMJ closed form (count+C1)" if grouped
                                                                    else "CPU_HJ (count+C1+C2)") +
                                         " restated in C, one window per thread (INTER)",
                               "seconds": secs,
                               "total_matches_equal_gpu": all(cpu_count == out[k]["total_matches"] for k in algs)}
    return out


def run_batch(a, np, torch, sjb, thetas, algs, steps, warmup):
    res = {}
    for th in thetas:
        res[f"theta_{th}"] = run_batch_theta(a, np, torch, sjb, th, algs, steps, warmup)
    return res


# ---------------------------------------------------------------------------
# learned vs radix partitioning
# ---------------------------------------------------------------------------
def learned_case(a, first, nw, torch, sjb, np, steps, warmup):
    """Fit on windows first-4..first-1, time windows first..first+nw-1 for
    HJ and SMJ, radix (plain) vs learned partitioning, both batched."""
    W = 1 << a.log_window
    train = np.concatenate([sjb.generate_mix(W, w, seed=1000 + w) for w in range(first - 4, first)])
    R = np.concatenate([sjb.generate_mix(W, w, seed=1000 + w) for w in range(first, first + nw)])
    S = np.concatenate([sjb.generate_mix(W, w, seed=5000 + w) for w in range(first, first + nw)])
    d_train = torch.from_numpy(train.view(np.uint8)).to("cuda")
    dR = torch.from_numpy(R.view(np.uint8)).to("cuda")
    dS = torch.from_numpy(S.view(np.uint8)).to("cuda")
    off = np.arange(nw + 1, dtype=np.uint64) * np.uint64(W)
    ctx = sjb.Context(device=0, max_window=W, distinct_keys=1 << 20)
    res = {}
    for name in ("HJ", "SMJ"):
        alg = sjb.ALG[name]
        out = {}
        # radix: hash (HJ) / radix-range (SMJ) partitions; radix_peel: the
        # same plus the batch's own sampled heavy keys aggregated apart (the
        # model-free baseline); learned: the fitted RMI (HJ: hash partitions
        # and the model's heavy keys; SMJ: CDF ranges and heavy keys)
        for mode in ("radix", "radix_peel", "learned"):
            f = sjb.F_GROUPED | {"radix": 0, "radix_peel": sjb.F_PEEL, "learned": sjb.F_LEARNED}[mode]
            if mode == "learned":
                fit = ctx.rmi_fit(d_train, len(train), root=sjb.RMI_MLP, steps=400, lr=0.01)
                out["fit"] = {k: fit.as_dict()[k] for k in ("fit_ms", "loss_init", "loss_final", "n_heavy",
                                                          "fallback")}
            else:
                ctx.rmi_clear()
            rep = ctx.partition_report(alg, f, dR, W, dS, W)
            best, tried = ctx.batch_autotune(alg, f, dR, off, dS, off, reps=2)
            d_res = torch.zeros(nw * 32, dtype=torch.uint8, device="cuda")
            t_plan = time.perf_counter()
            plan = ctx.batch_plan(alg, f, dR, off, dS, off, d_res)
            plan_ms = 1e3 * (time.perf_counter() - t_plan)
            pinfo = plan.info()
            ms, wall = time_plan(ctx, plan, steps, warmup, torch)
            r = d_res.cpu().numpy().view(sjb.WRES_DTYPE).copy()
            plan.destroy()
            med = float(np.median(ms))
            out[mode] = {"ms_per_step": med, "tuples_per_s": 2 * W * nw / (med / 1e3),
                         "p99_ms": float(np.percentile(ms, 99)),
                         "roofline_frac": 16 * 2 * W * nw / (med / 1e3) / HBM_PEAK,
                         "partition": {k: rep.as_dict()[k] for k in ("kind", "imbalance", "fixup_frac",
                                                                    "max_build", "heavy_left")},
                         "tune": best.as_dict(),
                         "plan_partitioning": ["hash", "radix ranges", "learned CDF"][pinfo.part_kind],
                         "drift_fallback": bool(pinfo.drift_fallback),
                         "drift_refit": bool(pinfo.drift_refit),
                         "heavy_keys_peeled": int(pinfo.n_heavy),
                         "plan_build_ms": plan_ms,
                         "results": [(int(x["count"]), int(x["checksum1"])) for x in r]}
        if not (out["radix"]["results"] == out["radix_peel"]["results"] == out["learned"]["results"]):
            raise SystemExit(f"{name}: learned, radix and radix+peeling results differ")
        out["speedup"] = out["radix"]["ms_per_step"] / out["learned"]["ms_per_step"]
        out["speedup_vs_radix_peel"] = out["radix_peel"]["ms_per_step"] / out["learned"]["ms_per_step"]
        out["peeling_alone_speedup"] = out["radix"]["ms_per_step"] / out["radix_peel"]["ms_per_step"]
        res[name] = out
    ctx.close()
    del d_train, dR, dS
    torch.cuda.empty_cache()
    # oracle: windows 0 and nw-1 (closed-form CPU_SMJ); CPU baseline: all
    # nw windows, one per thread
    orc = oracle_mod()
    results = res["HJ"]["radix"]["results"]
    fn = lambda w: orc.lazy_smj(R[w * W:(w + 1) * W], S[w * W:(w + 1) * W])[0]
    secs, cpu_res = cpu_windows_pool(fn, range(nw))
    ok = all((int(cpu_res[w].count), int(cpu_res[w].c1)) == tuple(results[w]) for w in range(nw))
    for name in res:
        for m in ("radix", "radix_peel", "learned"):
            res[name][m].pop("results")
    res["oracle_check"] = {"windows": nw, "equal": bool(ok)}
    res["cpu_baseline"] = {"value": 2 * W * nw / secs, "unit": "tuples/s", "cores": cpu_threads(), "kind": "port",
                           "sample": f"the same {nw} windows, CPU_SMJ closed form (count+C1) restated in C, "
                                     "one window per thread (INTER)", "seconds": secs}
    return res


def run_learned(a, np, torch, sjb, steps, warmup):
    """Learned vs plain radix partitioning on drifting keys: 2^22-tuple windows
    per side, keys 70% 8-component Gaussian mixture over u32 (means shift 1% of
    the range every 64 windows) + 30% Zipf(1.2).  The RMI (2x32 MLP root, 1024
    leaves, up to 32 heavy keys) is fitted on the 4 windows before the timed
    batch of 8 windows.  Same drift epoch: windows 68..75 (training 64..67);
    right after a drift step: windows 64..71, trained on 60..63 (the plan
    finds the model stale on its own key sample and refits it: drift_refit).
    Count + C1 (grouped).  A step is one replay of the batch plan."""
    nw = 8
    t0 = time.time()
    typical = learned_case(a, 68, nw, torch, sjb, np, steps, warmup)
    boundary = learned_case(a, 64, nw, torch, sjb, np, steps, warmup)
    def pick(d, key):
        return {k: v[key] for k, v in d.items() if isinstance(v, dict) and key in v}

    log(f"learned: {time.time() - t0:.1f}s; vs radix same epoch {pick(typical, 'speedup')} after drift "
        f"{pick(boundary, 'speedup')}; vs radix+peeling {pick(typical, 'speedup_vs_radix_peel')} / "
        f"{pick(boundary, 'speedup_vs_radix_peel')}")
    return {"workload": f"learned-partition HJ/SMJ, {nw} windows of 2^{a.log_window}/side per step, RMI "
                        "(2x32 MLP root, 1024 leaves) fitted on the 4 previous windows; count+C1 (grouped)",
            "speedup_vs_radix": pick(typical, "speedup"),
            "speedup_vs_radix_after_drift_step": pick(boundary, "speedup"),
            "speedup_vs_radix_peel": pick(typical, "speedup_vs_radix_peel"),
            "speedup_vs_radix_peel_after_drift_step": pick(boundary, "speedup_vs_radix_peel"),
            "note": "HJ keeps hash partitions: its learned gain is the model's heavy-key peeling, which "
                    "radix_peel (heavy keys of the batch's own sample, no model) also gets; the CDF model "
                    "itself acts on SMJ's range partitions",
            "same_epoch": typical, "after_drift_step": boundary}


# ---------------------------------------------------------------------------
# eager sliding window with a serving loop
# ---------------------------------------------------------------------------
def run_slide(a, np, torch, sjb):
    """Left and right batches arrive together as events, sizes Poisson(2^16)
    per side, keys Zipf(1.0) over [0, 2^26); the window is the last 2^22
    tuples of each side and every event's count + C1 is returned to the
    host.  Arrivals form a Poisson process at lambda input tuples/s, swept
    0.5-4 G/s, plus a back-to-back run (all events arrived at t = 0) for
    the saturated rate.  Latency = result host-visible minus arrival, over
    the events after the window has filled.  State is admitted under a
    64 GB HBM budget."""
    W, mean, n_events, fill = 1 << 22, 1 << 16, 1024 + 64, 64
    rng = np.random.default_rng(7)
    nl, nr = rng.poisson(mean, n_events), rng.poisson(mean, n_events)
    lo = np.concatenate([[0], np.cumsum(nl)]).astype(np.uint64)
    ro = np.concatenate([[0], np.cumsum(nr)]).astype(np.uint64)
    t0 = time.time()
    L = sjb.generate(int(lo[-1]), sjb.DIST_ZIPF, 1 << 26, 1.0, seed=123)
    R = sjb.generate(int(ro[-1]), sjb.DIST_ZIPF, 1 << 26, 1.0, seed=456)
    Lp = torch.from_numpy(L.view(np.uint8)).pin_memory()
    Rp = torch.from_numpy(R.view(np.uint8)).pin_memory()
    log(f"slide: {n_events} events, {int(lo[-1] + ro[-1])} tuples, generated in {time.time() - t0:.1f}s")
    ctx = sjb.Context(device=0, max_window=1 << 16, hbm_budget=64 << 30)
    sl = ctx.slide(W, 2 * mean)
    sizes = (nl + nr).astype(np.float64)
    rows, first = [], None
    for lam in (None, 0.5e9, 1e9, 2e9, 4e9):
        if lam is None:
            arrival = np.zeros(n_events)
        else:
            gaps = rng.exponential(sizes / lam)
            arrival = np.concatenate([[0.0], np.cumsum(gaps[:-1])])
        sl.reset()
        res, lat, el = sl.serve(Lp, lo, Rp, ro, arrival)
        if first is None:
            first = res.copy()
        elif not np.array_equal(first, res):
            raise SystemExit("sliding results differ between runs")
        m = lat[fill:] * 1e3
        rows.append({"offered_tuples_per_s": lam, "achieved_tuples_per_s": float(sizes.sum() / el),
                     "p50_ms": float(np.percentile(m, 50)), "p99_ms": float(np.percentile(m, 99)),
                     "max_ms": float(m.max()), "elapsed_s": el})
        log(f"slide lambda={lam}: {rows[-1]}")
    info = sl.info().as_dict()
    sl.destroy()
    ctx.close()
    sat = rows[0]
    # algorithmic bytes: each incoming tuple is read once (16 B) and leaves
    # the window once (16 B read again); state stays resident
    alg_bytes = 2 * 16 * float(sizes.sum())
    out = {"workload": "eager sliding window 2^22 tuples/side, events of Poisson(2^16) tuples per side, "
                       "Zipf(1.0) keys over [0, 2^26), host-pinned batches, count+C1 per event",
           "tuples_per_s": sat["achieved_tuples_per_s"], "ms_per_step": 1e3 * sat["elapsed_s"] / n_events,
           "steps": n_events, "latency": rows, "state": info,
           "roofline": {"bound": "hbm", "kernel": "event pipeline (H2D + four k_slide_step + snapshot)",
                        "achieved": alg_bytes / sat["elapsed_s"] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": alg_bytes / sat["elapsed_s"] / HBM_PEAK, "traffic": None,
                        "note": "bytes = 16 B per arriving + 16 B per expiring tuple; random table "
                                "accesses, not streaming, bound the step kernels"}}
    if not a.no_cpu_baseline:
        orc = oracle_mod()
        n_cpu = 256
        secs, cpu_res = orc.bench_slide(L, lo[1:n_cpu + 1], R, ro[1:n_cpu + 1], W)
        ok = cpu_res == [(int(x["count"]), int(x["checksum1"])) for x in first[:n_cpu]]
        tup = float(lo[n_cpu] + ro[n_cpu])
        out["cpu_baseline"] = {"value": tup / secs, "unit": "tuples/s", "cores": 1, "kind": "port",
                               "sample": f"events 0..{n_cpu - 1}: the same incremental count+C1 method with one "
                                         "hash map per side in C, one thread (the reference has no sliding "
                                         "window)", "seconds": secs, "results_match_gpu": bool(ok)}
        out["oracle_check"] = {"events": n_cpu, "equal": bool(ok)}
    return out


# ---------------------------------------------------------------------------
# eager per-batch joins (only with --workload eager)
# ---------------------------------------------------------------------------
def run_eager(a):
    """Eager (per-batch) joins on host batches, as the eager executors
    receive them: each batch is copied to the device, probes the other
    side's resident window state and is appended to its own side; the
    window (W tuples per side) closes when both sides are full.  Keys
    uniform in [0, 2^20), batches in host-pinned memory.  The
    persistent-kernel HJ serves batches from a resident grid (no launches
    per batch).  Per configuration: input tuples/s over all batches and
    p50/p99 of the per-batch call latency (host call to result, the H2D
    copy included).  Count + C1 + C2 per batch."""
    import numpy as np
    import torch
    import sjb

    cases = [("HJ", 1 << 20, 4, 1 << 14), ("PK_HJ", 1 << 20, 4, 1 << 14), ("NLJ", 1 << 17, 4, 1 << 14),
             ("HJ_small_batches", 1 << 16, 3, 1 << 10), ("PK_HJ_small_batches", 1 << 16, 3, 1 << 10)]
    rows = {}
    for name, W, nwin, B in cases:
        nb = W // B
        R = sjb.generate(W * nwin, sjb.DIST_UNIFORM, 1 << 20, 0.0, seed=123)
        S = sjb.generate(W * nwin, sjb.DIST_UNIFORM, 1 << 20, 0.0, seed=456)
        R = torch.from_numpy(R.view(np.uint8)).pin_memory().numpy().view(sjb.TUPLE_DTYPE)
        S = torch.from_numpy(S.view(np.uint8)).pin_memory().numpy().view(sjb.TUPLE_DTYPE)
        ctx = sjb.Context(device=0, max_window=W, distinct_keys=1 << 20, batch_size=B)
        pk = name.startswith("PK_HJ")
        alg = sjb.ALG["NLJ" if name == "NLJ" else "HJ"]
        lat, matches = [], 0

        def run(w, timed):
            nonlocal matches
            for b in range(nb):
                for is_left, src in ((1, R), (0, S)):
                    batch = src[(w * nb + b) * B:(w * nb + b + 1) * B]
                    t0 = time.perf_counter()
                    if pk:
                        rc, st = ctx.pk_execute(batch, is_left)
                    else:
                        rc, st = ctx.eager_execute(alg, sjb.OUT["NoOutput"], sjb.F_C2, batch, is_left)
                    if rc != 0:
                        raise SystemExit(f"eager {name}: rc {rc}")
                    if timed:
                        lat.append(time.perf_counter() - t0)
                        matches += st.result_count
            if pk:
                ctx.pk_clear()
            else:
                ctx.eager_clear()

        run(0, False)  # warm-up window
        t_start = time.perf_counter()
        for w in range(1, nwin):
            run(w, True)
        el = time.perf_counter() - t_start
        if pk:
            ctx.pk_stop()
        ctx.close()
        lat_ms = np.array(lat) * 1e3
        rows[name] = {"window_per_side": W, "batch": B, "windows_timed": nwin - 1,
                      "tuples_per_s": 2 * W * (nwin - 1) / el,
                      "p50_batch_ms": float(np.percentile(lat_ms, 50)),
                      "p99_batch_ms": float(np.percentile(lat_ms, 99)), "matches": int(matches)}
        log(f"eager {name}: {rows[name]}")
    return rows


def bench_eager(a):
    """--workload eager: the eager cases of run_eager as one JSON line (value:
    HJ with 2^14-tuple batches), with the oracle's eager hash join as the
    CPU baseline."""
    import sjb

    rows = run_eager(a)
    hj = rows["HJ"]
    out = {
        "metric": "eager per-batch join: input tuples/s and p50/p99 per-batch latency (host batches)",
        "value": hj["tuples_per_s"], "unit": "tuples/s", "n_gpus": 1,
        "steps": hj["windows_timed"], "warmup": 1,
        "ms_per_step": 1e3 * 2 * hj["window_per_side"] / hj["tuples_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 keys / u64 sums", "data": "synthetic (splitmix64 generator, uniform keys in [0, 2^20))",
        "config": {"workload": "eager tumbling windows, 2^14-tuple host-pinned batches alternating sides; HJ and "
                               "persistent-kernel HJ at 2^20 tuples/side, NLJ at 2^17; also HJ and "
                               "persistent-kernel HJ with 2^10-tuple batches at 2^16 tuples/side",
                   "checksums": "count+C1+C2", "parallelism": "single GPU"},
        "cases": rows,
    }
    if not a.no_cpu_baseline:
        orc = oracle_mod()
        W = 1 << 20
        R = sjb.generate(W, sjb.DIST_UNIFORM, 1 << 20, 0.0, seed=123)
        S = sjb.generate(W, sjb.DIST_UNIFORM, 1 << 20, 0.0, seed=456)
        e = orc.Eager(1 << 20, W)
        B = 1 << 14
        t0 = time.perf_counter()
        for b in range(W // B):
            e.execute(R[b * B:(b + 1) * B], True)
            e.execute(S[b * B:(b + 1) * B], False)
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 2 * W / secs, "unit": "tuples/s", "cores": 1, "kind": "port",
                               "sample": "one window of 2^20/side in 2^14-tuple batches, eager CPU_HJ "
                                         "restated in C, one thread", "seconds": secs}
    print(json.dumps(out))


def single_line(extra, metric, value, unit, a, workload, dtype="u32 keys / u64 sums"):
    """A stand-alone JSON line for one extra workload (--workload X)."""
    line = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": a.extra_steps,
            "warmup": a.warmup, "ms_per_step": extra.get("ms_per_step"), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic (splitmix64 generator)",
            "config": {"workload": workload, "parallelism": "single GPU"}}
    line.update({k: v for k, v in extra.items() if k not in line})
    return line


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the multi-GPU sharding is "
              "described in DESIGN.md and not measured.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "eager":
        return bench_eager(a)
    import numpy as np
    import torch
    import sjb

    thetas = [a.theta] if a.theta is not None else [0.0, 0.8, 1.2]
    if a.workload == "batch":
        res = run_batch(a, np, torch, sjb, thetas, [a.alg], a.steps, a.warmup)
        r = res[f"theta_{thetas[0]}"][a.alg]
        out = single_line(dict(r, detail=res), "joined input tuples/s + %HBM roofline, batched uneven windows",
                          r["tuples_per_s"], "tuples/s", a,
                          f"batched lazy {a.alg}: {a.n_windows} windows/launch, sizes LogNormal(2^14, 1) "
                          f"clipped [2^8, 2^20] per side, keys Zipf({thetas[0]}) over [0, 2^28)")
        out["steps"] = a.steps
        print(json.dumps(out))
        return
    if a.workload == "learned":
        r = run_learned(a, np, torch, sjb, a.steps, a.warmup)
        hj = r["same_epoch"]["HJ"]["learned"]
        out = single_line(r, "joined input tuples/s, learned+batched vs plain radix partitioning (drifting GMM + "
                             "Zipf 1.2)", hj["tuples_per_s"], "tuples/s", a, r["workload"])
        out["ms_per_step"], out["steps"] = hj["ms_per_step"], a.steps
        out["roofline"] = {"bound": "hbm", "kernel": "batch pipeline (learned HJ)",
                           "achieved": hj["roofline_frac"] * HBM_PEAK / 1e9, "peak": HBM_PEAK / 1e9,
                           "unit": "GB/s", "frac": hj["roofline_frac"], "traffic": None}
        print(json.dumps(out))
        return
    if a.workload == "slide":
        r = run_slide(a, np, torch, sjb)
        out = single_line(r, "sliding-window eager join: input tuples/s and p50/p99 per-event latency",
                          r["tuples_per_s"], "tuples/s", a, r["workload"])
        out["steps"] = r["steps"]
        print(json.dumps(out))
        return
    if a.workload == "nlj":
        print(json.dumps(run_nlj(a, np, torch, sjb, a.steps, a.warmup)))
        return
    if a.workload == "cfg1":
        r = run_cfg1(a, np, torch, sjb)
        out = single_line(r, "joined input tuples/s (cfg1: 2^16-tuple windows, keys U[0,2^16))",
                          r["gpu"]["tuples_per_s"], "tuples/s", a, r["workload"])
        out["ms_per_step"] = r["gpu"]["ms_per_step"]
        print(json.dumps(out))
        return

    t_all = time.time()
    out, R, S = run_window(a, np, torch, sjb)
    del R, S
    log(f"headline: {out['value'] / 1e9:.2f} G tuples/s, window p50 {out['window_ms']['p50']:.3f} ms "
        f"({time.time() - t_all:.0f}s)")
    if a.workload == "window":
        print(json.dumps(out))
        return
    extra, timing = {}, {}
    legs = [("cfg1", lambda: run_cfg1(a, np, torch, sjb)),
            ("nlj", lambda: run_nlj(a, np, torch, sjb, a.extra_steps, 1)),
            ("batch", lambda: run_batch(a, np, torch, sjb, thetas, ["HJ", "SMJ"], a.extra_steps, 1)),
            ("learned", lambda: run_learned(a, np, torch, sjb, a.extra_steps, 1)),
            ("slide", lambda: run_slide(a, np, torch, sjb)),
            ("eager", lambda: run_eager(a))]
    for name, fn in legs:
        t0 = time.time()
        try:
            extra[name] = fn()
        except SystemExit:
            raise
        except Exception as e:  # a failed extra leg is reported, the headline stands
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
        timing[name] = round(time.time() - t0, 1)
        torch.cuda.empty_cache()
        log(f"extra {name} done in {timing[name]}s")
    extra["wall_s"] = timing
    out["extra"] = extra
    print(json.dumps(out))


if __name__ == "__main__":
    main()
