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


M64 = (1 << 64) - 1


def splitmix64(j: int) -> int:
    """table[j] of the gather workload (first splitmix64 output with state j)."""
    z = (j + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


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


def bench_gather(B, torch, dev, steps, warmup, stream, n_log2=28, t_log2=28, tensors=None, cfg=None):
    n, tlen = 1 << n_log2, 1 << t_log2
    if tensors is None:
        table = torch.empty(tlen, dtype=torch.int64, device=dev)
        idx = torch.empty(n, dtype=torch.int64, device=dev)
        tgt = torch.empty(n, dtype=torch.int64, device=dev)
        B.fill_splitmix(table, stream)
        B.gen_index(idx, tlen, dist=0, seed=2, stream=stream)
    else:
        table, idx, tgt = tensors
    cfg = cfg or B.gather_plan(n, tlen)

    def step():
        B.gather(tgt, idx, table, cfg=cfg, stream=stream, sync=False)

    total, per = time_steps(step, steps, warmup, stream)
    assert B.sync_errors(stream) == 0
    # size-independent check on a sample: tgt[i] == splitmix64(idx[i])
    pos = torch.randint(0, n, (4096,), device=dev)
    iv = idx[pos].cpu().tolist()
    tv = tgt[pos].cpu().tolist()
    for a, b in zip(iv, tv):
        assert (b & M64) == splitmix64(a), "gather mismatch"
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


def time_calls(fn, reps, stream):
    """Median seconds of fn() over reps runs (one warm-up), HIP events on stream."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    return median(ts)


def check_mix_sample(mix, runs, k=None):
    """Exactness of every instance of a mix after `runs` applications, on the
    GPU: gathers against splitmix64 of the index (restated with torch integer
    arithmetic), histograms against torch.bincount.  k: only every
    (count / k)-th instance."""
    import torch
    sys.path.insert(0, str(ROOT / "tests"))
    from torch_check import splitmix64 as sm64
    step = max(1, len(mix.params) // k) if k else 1
    for i in range(0, len(mix.params), step):
        p = mix.params[i]
        tv, iv, gv = mix.views[i]
        if p.kind == 1:
            assert torch.equal(gv, sm64(iv)), ("gather", i)
        else:
            want = torch.bincount(iv, minlength=p.table_len) * runs
            assert torch.equal(tv.to(torch.int64), want), ("histo", i)


def bench_batch(B, torch, dev, stream, seed, count, reps, min_ops, learned):
    """The batched uneven mix: plain per-instance launches vs the persistent
    graph with exhaustive-tuned configurations (and with learned ones)."""
    from bale_old_amd import mix as MX
    from bale_old_amd import selector as S
    params = MX.mix_params(count, seed)
    mix = MX.Mix(params, dev, stream)
    torch.cuda.synchronize()
    ops = mix.total_ops
    out = {"instances": count, "seed": seed, "ops": ops,
           "table_bytes": sum(p.table_len for p in params) * 8, "index_bytes": ops * 8}
    t_plain = time_calls(lambda: MX.launch_plain(mix.inst, count, stream), reps, stream)
    B.sync_errors(stream)
    out["plain_ms"] = t_plain * 1e3
    out["plain_GUPS"] = ops / t_plain / 1e9
    log(f"seed {seed}: plain launches {t_plain * 1e3:.1f} ms; tuning ...")
    mix.reset_counts()
    t0 = time.time()
    MX.tune(mix.inst, count, min_ops=min_ops, stream=stream, chunk=512, log=lambda m: log(f"seed {seed}: {m}"))
    torch.cuda.synchronize()
    out["tune_seconds"] = time.time() - t0
    tuned = mix.configs()
    strat = {}
    for c in tuned:
        key = f"{MX.STRATEGIES.get(c.strategy, c.strategy)}{'+hot' if c.hot else ''}{'+runs' if c.merge > 0 else ''}"
        strat[key] = strat.get(key, 0) + 1
    out["exhaustive_strategies"] = strat
    bt = MX.Batch(mix.inst, count, stream)
    out["batch_stats"] = bt.stats()
    t_ex = time_calls(lambda: bt.launch(stream), reps, stream)
    B.sync_errors(stream)
    check_mix_sample(mix, reps + 1)
    bt.close()
    out["exhaustive_ms"] = t_ex * 1e3
    out["exhaustive_GUPS"] = ops / t_ex / 1e9
    out["exhaustive_vs_plain"] = t_plain / t_ex
    # what batching alone buys and what tuning alone buys: the tuned choices as
    # one library call per instance; the library's default rule and DIRECT
    # everywhere, each in one graph like the tuned batch
    mix.reset_counts()
    log(f"seed {seed}: exhaustive batch {t_ex * 1e3:.2f} ms")
    t_each = time_calls(lambda: MX.launch_each(mix.inst, count, stream), reps, stream)
    B.sync_errors(stream)
    check_mix_sample(mix, reps + 1)
    out["tuned_unbatched_ms"] = t_each * 1e3
    out["tuned_unbatched_GUPS"] = ops / t_each / 1e9
    for name, strat in (("default_rule_graph", B.AUTO), ("direct_graph", B.DIRECT)):
        mix.reset_counts()
        mix.set_configs([B.config(strat) for _ in params])
        bg = MX.Batch(mix.inst, count, stream)
        t_g = time_calls(lambda: bg.launch(stream), reps, stream)
        B.sync_errors(stream)
        check_mix_sample(mix, reps + 1)
        bg.close()
        out[f"{name}_ms"] = t_g * 1e3
        out[f"{name}_GUPS"] = ops / t_g / 1e9
        out[f"exhaustive_vs_{name}"] = t_g / t_ex
    mix.set_configs(tuned)
    # how often tuning the same instances again picks the same candidate (the
    # labels' own noise, an upper bound for any selector's agreement)
    sub = list(range(0, count, max(1, count // 256)))
    sub_inst = (MX.Instance * len(sub))(*[mix.inst[i] for i in sub])
    log(f"seed {seed}: re-tuning {len(sub)} instances ...")
    MX.tune(sub_inst, len(sub), min_ops=min_ops, stream=stream, chunk=128, log=lambda m: log(f"seed {seed}: {m}"))
    torch.cuda.synchronize()
    again = [MX.BaleConfigCopy(sub_inst[k].cfg) for k in range(len(sub))]
    out["retune_agreement"] = {"instances": len(sub),
                               "all_heads": sum(S.labels_of(again[k]) == S.labels_of(tuned[i])
                                                for k, i in enumerate(sub)) / len(sub),
                               "per_head": [sum(S.labels_of(again[k])[h] == S.labels_of(tuned[i])[h]
                                                for k, i in enumerate(sub)) / len(sub) for h in range(3)]}
    if learned is not None:
        mix.reset_counts()
        t0 = time.time()
        S.select(mix.inst, count, learned, dev, stream)
        torch.cuda.synchronize()
        out["select_seconds"] = time.time() - t0
        got = mix.configs()
        agree = sum(1 for a, b in zip(tuned, got) if S.labels_of(a) == S.labels_of(b))
        out["learned_agrees_with_exhaustive"] = agree / count
        out["learned_head_agreement"] = [sum(S.labels_of(a)[h] == S.labels_of(b)[h]
                                             for a, b in zip(tuned, got)) / count for h in range(3)]
        bl = MX.Batch(mix.inst, count, stream)
        t_l = time_calls(lambda: bl.launch(stream), reps, stream)
        B.sync_errors(stream)
        check_mix_sample(mix, reps + 1)
        bl.close()
        out["learned_ms"] = t_l * 1e3
        out["learned_GUPS"] = ops / t_l / 1e9
        out["learned_vs_exhaustive"] = t_ex / t_l
        out["learned_vs_plain"] = t_plain / t_l
    del mix
    torch.cuda.empty_cache()
    return out


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


def cpu_share():
    """Threads for the CPU baselines: the CPU share this process may use (the
    cores it is bound to, at most OMP_NUM_THREADS when that is set, as on the
    GPU boxes), with nproc and the affinity recorded beside it."""
    aff = len(os.sched_getaffinity(0))
    share = aff
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        share = min(aff, int(env))
    return share, {"nproc": os.cpu_count(), "affinity_cpus": aff, "threads_used": share,
                   "limit": "OMP_NUM_THREADS" if share < aff else "affinity"}


def cpu_baseline(idx_t, table_t):
    """The reference's two models of the gather on the host cores, timed like
    the reference (per-PE wall seconds between barriers, averaged over PEs,
    libgetput/libgetput.upc:395-401): AGI (relaxed loads, apps/ig_src/ig_agi.upc:61-64)
    and aggregate-by-owner request/reply (apps/ig_src/ig_exstack.upc:69-94),
    at P = the CPU share and P = 1."""
    import numpy as np
    import oracle
    P, info = cpu_share()
    n = idx_t.numel()
    sample = min(n, 1 << 26)
    idx = idx_t[:sample].cpu().numpy()
    table = table_t.cpu().numpy()
    tgt = np.empty(sample, np.int64)
    s1 = min(n, 1 << 23)

    def timed(fn, m, threads):
        fn(idx[:m], table, tgt[:m], threads)  # warm-up
        t = sum(fn(idx[:m], table, tgt[:m], threads) for _ in range(3)) / 3
        assert np.array_equal(tgt[:m], table[idx[:m]])
        return m / t / 1e9

    models = {}
    for name, fn in (("agi", oracle.ig_agi_mt), ("owner", oracle.ig_owner_mt)):
        models[name] = {f"P{P}": timed(fn, sample, P), "P1": timed(fn, s1, 1)}
    best = max(models, key=lambda k: models[k][f"P{P}"])
    return {"value": models[best][f"P{P}"], "unit": "GUPS", "cores": P, "kind": "port",
            "model": best, "models_GUPS": models, "host": info,
            "sample": f"first 2^{sample.bit_length() - 1} of the 2^28 requests into the full 2^28-entry "
                      f"table (2 GiB) with {P} threads (2^{s1.bit_length() - 1} with 1), the oracle's "
                      f"AGI and owner models, mean of 3 after 1 warm-up; value = the faster at P={P}"}


def cpu_histo():
    """The reference's AGI histogram (relaxed atomic adds on shared counters,
    apps/histo_src/histo_agi.upc:61-66) and aggregate-by-owner one
    (apps/histo_src/histo_exstack.upc:64-76) on the host cores and on one core:
    uniform updates into 2^26 int64 counters, the bench's 2^28-update case
    sampled to 2^26 (2^24 for one thread)."""
    import numpy as np
    import oracle
    P, info = cpu_share()
    nbins = 1 << 26
    idx = oracle.gen_index(1 << 26, nbins, 0, 1.0, 0, 3)
    counts = np.zeros(nbins, np.int64)
    res = {"host": info}
    for name, fn in (("agi", oracle.histo_agi_mt), ("owner", oracle.histo_owner_mt)):
        for cores, n in ((P, idx.size), (1, 1 << 24)):
            fn(idx[:n], counts, cores)
            t = sum(fn(idx[:n], counts, cores) for _ in range(3)) / 3
            res[f"{name}_P{cores}"] = {"value": n / t / 1e9, "unit": "GUPS", "cores": cores, "kind": "port",
                                       "sample": f"2^{n.bit_length() - 1} uniform updates (seed 3) into 2^26 "
                                                 f"int64 counters, mean of 3 after 1 warm-up"}
    return res


def load_traffic(name: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    p = ROOT / "profiles" / f"{name}_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes:"):
            return int(line.split(":", 1)[1])
    return None


def bench_convey(B, torch, dev, stream, n=1 << 26, P=64, reps=5):
    """Device conveyor: push 2^26 8-byte items to 64 virtual PEs and deliver
    them (advance), wall time of push + advance with the GPU idle before."""
    from bale_old_amd.convey import Conveyor, DONE
    items = torch.empty(n, dtype=torch.int64, device=dev)
    B.fill_splitmix(items)
    pes = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(pes, P, seed=6, stream=stream)
    want = torch.bincount(pes, minlength=P).cpu().tolist()
    c = Conveyor(P, n, 8)
    c.begin()
    ts = []
    for rep in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert c.push(items, pes, stream)
        c.advance(False, stream)
        t1 = time.perf_counter()
        got = [c.pull_span(p)[1] for p in range(P)]
        assert got == want, "conveyor spans do not match the PE counts"
        if rep:
            ts.append(t1 - t0)
    assert c.advance(True, stream) == DONE
    c.free()
    t = median(ts)
    return {"n": n, "pes": P, "item_bytes": 8, "median_s": t, "Gitems_per_s": n / t / 1e9,
            "note": "push (device copy of items and PE numbers) + advance (count, offsets, "
                    "stable delivery, span readback)"}


def cpu_convey(n=1 << 24, P=64):
    """The oracle's FIFO delivery of the same kind of input, one host thread."""
    import numpy as np
    import oracle
    rng = np.random.default_rng(6)
    items = rng.integers(0, 1 << 62, size=n, dtype=np.int64)
    pes = rng.integers(0, P, size=n, dtype=np.int64)
    oracle.convey_deliver(items, pes, P)
    t0 = time.perf_counter()
    for _ in range(3):
        oracle.convey_deliver(items, pes, P)
    t = (time.perf_counter() - t0) / 3
    return {"n": n, "pes": P, "Gitems_per_s": n / t / 1e9, "cores": 1, "kind": "port"}


def bench_randperm(B, torch, dev, stream, N=1 << 26, reps=5, seed=101892):
    """rand_permp_agi at one PE on the GPU: 2^26 elements (about the reference
    driver's 10^6 per thread at 64 threads), seed 101892 (its PE-0 seed)."""
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm(perm, seed, stream)
    assert B.is_perm(perm, stream), "randperm output is not a permutation"
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        B.randperm(perm, seed, stream)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t = median(ts)
    darts = int(1.3862944 * N + 8.0 * N ** 0.5) + 4096  # darts generated
    return {"n": N, "seed": seed, "median_s": t, "Gelems_per_s": N / t / 1e9, "darts": darts,
            "note": "wall time of one call (darts sorted by 2^15-slot region, each slot's first dart "
                    "and the hits' ranks resolved in LDS, items placed region by region, compaction; "
                    "one 4-byte readback of the hit count)"}


def cpu_randperm(N=1 << 24, seed=101892):
    """The oracle's sequential dart loop (glibc lrand48), one host thread."""
    import oracle
    oracle.randperm_agi(1 << 16, seed)
    t0 = time.perf_counter()
    oracle.randperm_agi(N, seed)
    t = time.perf_counter() - t0
    return {"n": N, "Gelems_per_s": N / t / 1e9, "cores": 1, "kind": "port"}


def bench_spmat(B, torch, dev, stream, reps=3):
    """Sparse-matrix kernels on the GPU, wall time of one call each (inputs
    resident): transpose and permute of an Erdos-Renyi matrix of 2^24 rows
    (the transpose driver's density: p = 18 / n, both triangles), and
    toposort of the toposort driver's input at 2^20 rows (upper triangle with
    a full diagonal, 35 per row, rows and columns permuted)."""
    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return median(ts), r
    out = {}
    n = 1 << 24
    off, nz = B.spmat_gen_er(n, 18 / n, 0, 3, seed=101894)
    nnz = nz.numel()
    t, (toff, tnz) = timed(lambda: B.spmat_transpose(n, off, nz))
    ooff, onz = B.spmat_transpose(n, toff, tnz)
    assert bool((ooff == off).all()) and bool((onz == nz).all()), "transpose of the transpose differs"
    # 16 algorithmic bytes per nonzero: A's column read, the transpose's row written
    out["transpose"] = {"rows": n, "nnz": nnz, "median_s": t, "Gnnz_per_s": nnz / t / 1e9,
                        "alg_GBps": 16 * nnz / t / 1e9}
    del toff, tnz, ooff, onz
    rpi = torch.empty(n, dtype=torch.int64, device=dev)
    cpi = torch.empty(n, dtype=torch.int64, device=dev)
    B.randperm(rpi, 1230)
    B.randperm(cpi, 45)
    t, _ = timed(lambda: B.spmat_permute(n, off, nz, rpi, cpi))
    # one random 8-byte read of cperminv per nonzero (a 128 MiB table)
    out["permute"] = {"rows": n, "nnz": nnz, "median_s": t, "Gnnz_per_s": nnz / t / 1e9,
                      "random_reads_G_per_s": nnz / t / 1e9}
    del off, nz, rpi, cpi
    torch.cuda.empty_cache()
    n = 1 << 20
    off, nz = B.spmat_gen_er(n, 70 / (n - 1), 1, 2, seed=5)
    rpi = torch.empty(n, dtype=torch.int64, device=dev)
    cpi = torch.empty(n, dtype=torch.int64, device=dev)
    B.randperm(rpi, 1230)
    B.randperm(cpi, 45)
    moff, mnz = B.spmat_permute(n, off, nz, rpi, cpi)
    toff, tnz = B.spmat_transpose(n, moff, mnz)
    t, (rp, cp, lev) = timed(lambda: B.toposort(moff, mnz, toff, tnz))
    po, pn = B.spmat_permute(n, moff, mnz, rp, cp)
    assert B.spmat_triangular(po, pn, True) == (0, 0), "toposort result is not triangular"
    # one device atomic per nonzero (packed count and column sum); event keys
    # only for the rows that join the next level
    out["toposort"] = {"rows": n, "nnz": mnz.numel(), "levels": lev, "median_s": t,
                       "Grows_per_s": n / t / 1e9, "Gnnz_per_s": mnz.numel() / t / 1e9,
                       "device_atomics_G_per_s": mnz.numel() / t / 1e9}
    out["note"] = "wall time of one call with a stream sync; bit-exact vs the oracle in tests/"
    return out


def cpu_spmat():
    """The oracle's AGI loops at one PE on one host thread, smaller samples:
    transpose / permute of a 2^22-row matrix of the same density, toposort at
    2^20 rows."""
    import numpy as np
    import oracle
    res = {}
    n = 1 << 22
    off, nz = oracle.gen_er(n, 18 / n, 0, 3, 101894)
    t0 = time.perf_counter()
    oracle.transpose_agi(n, n, off, nz)
    t = time.perf_counter() - t0
    res["transpose"] = {"rows": n, "nnz": int(nz.size), "Gnnz_per_s": nz.size / t / 1e9, "cores": 1, "kind": "port"}
    rng = np.random.default_rng(1)
    rpi, cpi = rng.permutation(n).astype(np.int64), rng.permutation(n).astype(np.int64)
    t0 = time.perf_counter()
    oracle.permute_agi(off, nz, rpi, cpi)
    t = time.perf_counter() - t0
    res["permute"] = {"rows": n, "nnz": int(nz.size), "Gnnz_per_s": nz.size / t / 1e9, "cores": 1, "kind": "port"}
    n = 1 << 20
    off, nz = oracle.gen_er(n, 70 / (n - 1), 1, 2, 5)
    rpi, cpi = rng.permutation(n).astype(np.int64), rng.permutation(n).astype(np.int64)
    moff, mnz = oracle.permute_agi(off, nz, rpi, cpi)
    toff, tnz = oracle.transpose_agi(n, n, moff, mnz)
    t0 = time.perf_counter()
    oracle.toposort_agi(moff, mnz, toff, tnz)
    t = time.perf_counter() - t0
    res["toposort"] = {"rows": n, "nnz": int(mnz.size), "Grows_per_s": n / t / 1e9,
                       "Gnnz_per_s": mnz.size / t / 1e9, "cores": 1, "kind": "port"}
    return res


WORKLOADS = {
    "gather": "the planned gather (by table region) of 2^28 uniform requests into a 2^28-entry table",
    "gather_direct": "k_gather_direct over the same requests",
    "copy": "k_probe_copy of exactly 1 GiB (read 1 GiB, write 1 GiB): the counters' calibration",
    "histo": "the planned histogram of 2^28 uniform updates into 2^26 int64 bins",
    "zipf": "the same plan on 2^28 Zipf(1.1) updates (its sample switches the hot table on)",
    "small": "the planned histogram of 2^24 uniform updates into 2^20 int64 bins",
    "batch": "the held-out mix of 4096 instances (seed 5), learned choices, launched three times",
    "transpose": "three transposes of the Erdos-Renyi matrix of 2^24 rows (p = 18 / n, both triangles)",
    "randperm": "three random permutations of 2^26 elements (rand_permp_agi, seed 101892)",
    "permute": "three permutes of the Erdos-Renyi matrix of 2^24 rows (row and column permutations "
               "from rand_permp, seeds 1230 and 45)",
    "toposort": "three toposorts of the benchmark's 2^20-row input (upper-triangular Erdos-Renyi with "
                "a full diagonal, 35 per row, rows and columns permuted)",
}


def run_workloads(which) -> int:
    """One call of each named workload and nothing else, for hardware-counter
    runs (rocprofv3 --pmc ... -- python3 bench.py --workload gather copy)."""
    import torch
    import bale_old_amd as B
    dev = torch.device("cuda:0")
    B.check(B.lib().bale_hip_init(0))
    st = torch.cuda.current_stream()
    n = 1 << 28
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    if which & {"gather", "gather_direct", "copy"}:
        table = torch.empty(n, dtype=torch.int64, device=dev)
        B.fill_splitmix(table)
        tgt = torch.empty(n, dtype=torch.int64, device=dev)
        B.gen_index(idx, n, seed=2)
        if "gather" in which:
            B.gather(tgt, idx, table, cfg=B.gather_plan(n, n))
        if "gather_direct" in which:
            B.gather(tgt, idx, table, cfg=B.config(B.DIRECT, block=256, unroll=8))
        if "copy" in which:
            B.check(B.lib().bale_probe_copy(tgt.data_ptr(), table.data_ptr(), 1 << 30, st.cuda_stream))
        torch.cuda.synchronize()
        del table, tgt
    for name, nb, dist in (("histo", 1 << 26, 0), ("zipf", 1 << 26, 1)):
        if name in which:
            counts = torch.zeros(nb, dtype=torch.int64, device=dev)
            B.gen_index(idx, nb, dist=dist, s=1.1, seed=3)
            B.histo(idx, counts, cfg=B.histo_plan(n, nb))
            del counts
    if "small" in which:
        m, nb = 1 << 24, 1 << 20
        counts = torch.zeros(nb, dtype=torch.int64, device=dev)
        B.gen_index(idx[:m], nb, seed=1)
        B.histo(idx[:m], counts, cfg=B.histo_plan(m, nb))
        del counts
    if "transpose" in which:
        nr = 1 << 24
        off, nz = B.spmat_gen_er(nr, 18 / nr, 0, 3, seed=101894)
        for _ in range(3):
            toff, tnz = B.spmat_transpose(nr, off, nz)
        torch.cuda.synchronize()
        del off, nz, toff, tnz
    if "permute" in which:
        nr = 1 << 24
        off, nz = B.spmat_gen_er(nr, 18 / nr, 0, 3, seed=101894)
        rpi = torch.empty(nr, dtype=torch.int64, device=dev)
        cpi = torch.empty(nr, dtype=torch.int64, device=dev)
        B.randperm(rpi, 1230)
        B.randperm(cpi, 45)
        for _ in range(3):
            poff, pnz = B.spmat_permute(nr, off, nz, rpi, cpi)
        torch.cuda.synchronize()
        del off, nz, rpi, cpi, poff, pnz
    if "toposort" in which:
        nr = 1 << 20
        off, nz = B.spmat_gen_er(nr, 70 / (nr - 1), 1, 2, seed=5)
        rpi = torch.empty(nr, dtype=torch.int64, device=dev)
        cpi = torch.empty(nr, dtype=torch.int64, device=dev)
        B.randperm(rpi, 1230)
        B.randperm(cpi, 45)
        moff, mnz = B.spmat_permute(nr, off, nz, rpi, cpi)
        toff, tnz = B.spmat_transpose(nr, moff, mnz)
        for _ in range(3):
            B.toposort(moff, mnz, toff, tnz)
        torch.cuda.synchronize()
        del off, nz, rpi, cpi, moff, mnz, toff, tnz
    if "randperm" in which:
        perm = torch.empty(1 << 26, dtype=torch.int64, device=dev)
        for _ in range(3):
            B.randperm(perm, 101892)
        torch.cuda.synchronize()
        del perm
    if "batch" in which:
        del idx
        from bale_old_amd import mix as MX, selector as S
        mix = MX.Mix(MX.mix_params(4096, 5), dev, st)
        weights, _ = S.load(dev)
        S.select(mix.inst, 4096, weights, dev, st)
        b = MX.Batch(mix.inst, 4096, st)
        for _ in range(3):
            b.launch(st)
        torch.cuda.synchronize()
        b.close()
    torch.cuda.synchronize()
    log("workloads done: " + " ".join(sorted(which)))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="headline gather only")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched mixes")
    ap.add_argument("--batch-instances", type=int, default=4096)
    ap.add_argument("--workload", nargs="+", choices=sorted(WORKLOADS),
                    help="run only these workloads once each (for rocprofv3 counter runs) and exit: "
                         + "; ".join(f"{k}: {v}" for k, v in WORKLOADS.items()))
    args = ap.parse_args()
    if args.workload:
        return run_workloads(set(args.workload))
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the exchange between GPUs is designed, "
              "not built or measured; run with --gpus 1", file=sys.stderr)
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

    log("gather 2^28 x 2^28 (planned: by table region, sample-checked) ...")
    g = bench_gather(B, torch, dev, args.steps, args.warmup, stream)
    table, idx, tgt = g.pop("tensors")
    ms = g["total_s"] / args.steps * 1e3
    value = g["n"] * args.steps / g["total_s"] / 1e9
    kern_s = sum(g["per_step_s"]) / len(g["per_step_s"])
    alg_bytes = 24 * g["n"]  # 8 B index + 8 B table + 8 B tgt per request
    achieved = alg_bytes / kern_s / 1e9
    strategy = {B.DIRECT: "direct", B.BUCKET: "bucket"}.get(g["cfg"]["strategy"], str(g["cfg"]["strategy"]))
    log(f"gather: {value:.2f} GUPS, {ms:.3f} ms/step, {achieved:.0f} GB/s algorithmic ({strategy})")

    extra = {"gather_cfg": g["cfg"], "gather_median_step_ms": median(g["per_step_s"]) * 1e3}
    # the one-load-per-request kernel on the same tensors, for comparison
    gd = bench_gather(B, torch, dev, args.steps, args.warmup, stream, tensors=(table, idx, tgt),
                      cfg=B.config(B.DIRECT, block=256, unroll=8))
    gd.pop("tensors")
    extra["gather_direct"] = {"GUPS": gd["n"] * args.steps / gd["total_s"] / 1e9,
                              "ms_per_step": gd["total_s"] / args.steps * 1e3, "cfg": gd["cfg"]}
    log(f"gather direct: {extra['gather_direct']}")
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
            # algorithmic bytes: 8 B index + counter read-modify-write (16 B int64, 8 B int32)
            r["alg_GBps"] = r["GUPS"] * (24 if kw["dtype"] == torch.int64 else 16)
            histo[name] = r
            log(f"{name}: {r['GUPS']:.2f} GUPS cfg={r['cfg']}")
            torch.cuda.empty_cache()
        if not args.no_cpu:
            histo["cpu_baseline"] = cpu_histo()
            log(f"histo cpu: {histo['cpu_baseline']}")
        extra["histo"] = histo
        log("conveyor 2^26 x 8 B -> 64 PEs ...")
        extra["convey"] = bench_convey(B, torch, dev, stream)
        if not args.no_cpu:
            extra["convey"]["cpu_baseline"] = cpu_convey()
        log(f"conveyor: {extra['convey']}")
        log("randperm 2^26 ...")
        extra["randperm"] = bench_randperm(B, torch, dev, stream)
        if not args.no_cpu:
            extra["randperm"]["cpu_baseline"] = cpu_randperm()
        log(f"randperm: {extra['randperm']}")
        log("sparse matrices: transpose, permute 2^24 rows; toposort 2^20 rows ...")
        extra["spmat"] = bench_spmat(B, torch, dev, stream)
        if not args.no_cpu:
            extra["spmat"]["cpu_baseline"] = cpu_spmat()
        log(f"spmat: {extra['spmat']}")
        torch.cuda.empty_cache()

    if not args.no_batch:
        from bale_old_amd import selector as S
        learned = None
        if S.WEIGHTS.exists():
            learned, meta = S.load(dev)
            extra["selector_training"] = meta
        log("batched mix, seed 4 (plain vs exhaustive) ...")
        extra["batch_mix_seed4"] = bench_batch(B, torch, dev, stream, 4, args.batch_instances, 3,
                                               1 << 19, None)
        log(f"seed 4: {extra['batch_mix_seed4']}")
        log("held-out mix, seed 5 (learned vs exhaustive vs plain) ...")
        extra["batch_mix_seed5"] = bench_batch(B, torch, dev, stream, 5, args.batch_instances, 3,
                                               1 << 19, learned)
        log(f"seed 5: {extra['batch_mix_seed5']}")

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
                   "requests": g["n"], "table_len": g["table_len"], "strategy": strategy},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic("gather_bucket_2p28" if strategy == "bucket" else "gather_direct_2p28"),
                     "kernel": ("bucketed gather call: k_gsort + k_gtranspose + k_gregion + k_gunpermute "
                                "(+ the sample and the DIRECT launch it skips)" if strategy == "bucket"
                                else "k_gather_direct"),
                     "kernel_avg_ms": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
