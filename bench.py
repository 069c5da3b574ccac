#!/usr/bin/env python3
"""Benchmark of bale's index gather and histogram on one MI355X.

One step = one pass of the hot path over one batch of synthetic input already
resident in HBM.  The headline workload is the plain index gather: 2^28 int64
requests, uniform over a 2^28-entry int64 table with table[j] = splitmix64(j)
(indices from splitmix64 seed 2), i.e. 2 GiB table + 2 GiB index + 2 GiB tgt.
`value` is its rate in GUPS (10^9 requests per second).  The histogram
configurations are measured in the same run and reported under "extra".

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GUPS for histo & ig at 2^28 updates; learned-vs-plain batched speedup"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(fn, steps: int, warmup: int, stream):
    """Run fn() warmup times, then time `steps` calls one by one with HIP events
    on `stream`.  Returns (total seconds, list of per-step seconds)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t_all0 = torch.cuda.Event(enable_timing=True)
    t_all1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_all0.record(stream)
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    t_all1.record(stream)
    torch.cuda.synchronize()
    per = [a.elapsed_time(b) * 1e-3 for a, b in evs]
    return t_all0.elapsed_time(t_all1) * 1e-3, per


def median(xs):
    s = sorted(xs)
    return s[len(s) // 2]


def bench_gather(B, torch, dev, steps, warmup, stream, n_log2=28, t_log2=28):
    n, tlen = 1 << n_log2, 1 << t_log2
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    tgt = torch.empty(n, dtype=torch.int64, device=dev)
    B.fill_splitmix(table, stream)
    B.gen_index(idx, tlen, dist=0, seed=2, stream=stream)
    cfg = B.gather_plan(n, tlen)

    def step():
        B.gather(tgt, idx, table, cfg=cfg, stream=stream, sync=False)

    total, per = time_steps(step, steps, warmup, stream)
    assert B.sync_errors(stream) == 0
    # size-independent check on a sample: tgt[i] == splitmix64(idx[i])
    import oracle
    pos = torch.randint(0, n, (4096,), device=dev)
    iv = idx[pos].cpu().tolist()
    tv = tgt[pos].cpu().tolist()
    for a, b in zip(iv, tv):
        assert (b & 0xFFFFFFFFFFFFFFFF) == oracle.splitmix64(a), "gather mismatch"
    return {"n": n, "table_len": tlen, "total_s": total, "per_step_s": per, "cfg": cfg.as_dict(),
            "tensors": (table, idx, tgt)}


def bench_histo(B, torch, dev, steps, warmup, stream, n, nbins, dtype, dist, s, run_mean, seed,
                cfg=None):
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(idx, nbins, dist=dist, s=s, run_mean=run_mean, seed=seed, stream=stream)
    counts = torch.zeros(nbins, dtype=dtype, device=dev)
    cfg = cfg or B.histo_plan(n, nbins, B.I64 if dtype == torch.int64 else B.I32)

    def step():
        B.histo(idx, counts, cfg=cfg, stream=stream, sync=False)

    total, per = time_steps(step, steps, warmup, stream)
    assert B.sync_errors(stream) == 0
    # every step adds n updates
    got = int(counts.sum(dtype=torch.int64).item())
    assert got == n * (steps + warmup), (got, n * (steps + warmup))
    del idx, counts
    return {"n": n, "nbins": nbins, "total_s": total, "median_s": median(per), "cfg": cfg.as_dict()}


def probes(B, torch, dev, stream, table):
    """Streaming copy rate and random 8-byte read rate on this GPU."""
    nbytes = 1 << 31
    src = torch.empty(nbytes // 8, dtype=torch.int64, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1)
    lib = B.lib()
    sp = stream.cuda_stream

    def cp():
        B.check(lib.bale_probe_copy(dst.data_ptr(), src.data_ptr(), nbytes, sp))

    _, per = time_steps(cp, 10, 2, stream)
    copy_gbs = 2 * nbytes / median(per) / 1e9
    del src, dst
    out = torch.zeros(1, dtype=torch.int64, device=dev)
    nreads = 1 << 28

    def rr():
        B.check(lib.bale_probe_random_read(table.data_ptr(), table.numel(), nreads, out.data_ptr(), sp))

    _, per = time_steps(rr, 10, 2, stream)
    rand_gups = nreads / median(per) / 1e9
    return {"copy_GBps": copy_gbs, "random_read_GUPS": rand_gups,
            "random_read_table_bytes": table.numel() * 8}


def cpu_baseline(idx_t, table_t, max_threads=16):
    """The oracle's multi-threaded AGI gather (relaxed loads) on the host cores,
    timed like the reference (per-thread wall seconds between barriers,
    averaged over threads, libgetput/libgetput.upc:395-401)."""
    import numpy as np
    import oracle
    threads = max(1, min(max_threads, len(os.sched_getaffinity(0))))
    n = idx_t.numel()
    sample = min(n, 1 << 26)
    idx = idx_t[:sample].cpu().numpy()
    table = table_t.cpu().numpy()
    tgt = np.empty(sample, np.int64)
    oracle.ig_agi_mt(idx, table, tgt, threads)  # warm-up
    secs = []
    for _ in range(3):
        secs.append(oracle.ig_agi_mt(idx, table, tgt, threads))
    t = sum(secs) / len(secs)
    return {"value": sample / t / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
            "sample": f"first 2^{sample.bit_length() - 1} of the 2^28 requests into the full 2^28-entry "
                      f"table (2 GiB), oracle or_ig_agi_mt with {threads} threads, mean of 3 after 1 warm-up"}


def load_traffic(name: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("traffic_bytes")
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="headline gather only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU exchange is a design (see DESIGN.md), "
              "not a measured path; run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    import bale_old_amd as B
    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    B.check(B.lib().bale_hip_init(0))
    stream = torch.cuda.current_stream()
    cus, lds = B.device_info()
    log(f"device: {torch.cuda.get_device_name(0)} cus={cus} lds/wg={lds}")

    log("gather 2^28 x 2^28 ...")
    g = bench_gather(B, torch, dev, args.steps, args.warmup, stream)
    table, idx, tgt = g.pop("tensors")
    ms = g["total_s"] / args.steps * 1e3
    value = g["n"] * args.steps / g["total_s"] / 1e9
    kern_s = sum(g["per_step_s"]) / len(g["per_step_s"])
    alg_bytes = 24 * g["n"]  # 8 B index + 8 B table + 8 B tgt per request
    achieved = alg_bytes / kern_s / 1e9
    log(f"gather: {value:.2f} GUPS, {ms:.3f} ms/step, {achieved:.0f} GB/s algorithmic")

    extra = {"gather_cfg": g["cfg"], "gather_median_step_ms": median(g["per_step_s"]) * 1e3}
    if not args.no_extra:
        log("probes ...")
        pr = probes(B, torch, dev, stream, table)
        extra["probes"] = pr
        extra["gather_vs_random_read_ceiling"] = value / pr["random_read_GUPS"]
        log(f"probes: {pr}")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(idx, table)
        log(f"cpu: {cpu}")
    del table, idx, tgt
    torch.cuda.empty_cache()

    if not args.no_extra:
        hsteps = max(3, args.steps // 2)
        histo = {}
        for name, kw in [
            ("histo_plain_2^24_into_2^20_i64", dict(n=1 << 24, nbins=1 << 20, dtype=torch.int64, dist=0, s=1.0, run_mean=0, seed=1)),
            ("histo_uniform_2^28_into_2^26_i64", dict(n=1 << 28, nbins=1 << 26, dtype=torch.int64, dist=0, s=1.0, run_mean=0, seed=3)),
            ("histo_zipf1.1_2^28_into_2^26_i64", dict(n=1 << 28, nbins=1 << 26, dtype=torch.int64, dist=1, s=1.1, run_mean=0, seed=3)),
            ("histo_zipf1.1_2^28_into_2^26_i32", dict(n=1 << 28, nbins=1 << 26, dtype=torch.int32, dist=1, s=1.1, run_mean=0, seed=3)),
            ("histo_zipf1.1_runs64_2^28_into_2^26_i64", dict(n=1 << 28, nbins=1 << 26, dtype=torch.int64, dist=1, s=1.1, run_mean=64, seed=3)),
        ]:
            log(f"{name} ...")
            r = bench_histo(B, torch, dev, hsteps, 2, stream, **kw)
            r["GUPS"] = r["n"] / r["median_s"] / 1e9
            histo[name] = r
            log(f"{name}: {r['GUPS']:.2f} GUPS cfg={r['cfg']}")
            torch.cuda.empty_cache()
        extra["histo"] = histo

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded splitmix64 streams generated on the GPU)",
        "config": {"workload": "index gather: 2^28 int64 requests, uniform (splitmix64 seed 2) over a "
                               "2^28-entry int64 table with table[j]=splitmix64(j); inputs resident in HBM",
                   "requests": g["n"], "table_len": g["table_len"], "strategy": "direct"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("gather_direct_2p28"),
                     "kernel": "k_gather_direct", "kernel_avg_ms": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
