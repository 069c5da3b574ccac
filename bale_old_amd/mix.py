"""The batched uneven mix: many histogram / gather instances of log-uniform
sizes, generated on the GPU and handed to libbale_hip.so as one batch.

Instance t of a mix with seed s draws six values from the splitmix64 stream
seeded with the t-th output of stream s:
  d0 -> kind (top bit: 0 histogram, 1 gather)
  d1 -> table_len = floor(2^(lo_tab + (hi_tab-lo_tab)*u1))
  d2 -> nops      = floor(2^(lo_ops + (hi_ops-lo_ops)*u2))
  d3 -> dist      = reduce(d3, 3): 0 uniform, 1 Zipf(s), 2 clustered runs (mean 64)
  d4 -> s         = 0.8 + 0.6*u4
  d5 -> index seed
with u = (d >> 11) * 2^-53.  Gather tables hold table[j] = splitmix64(j);
histogram counts start at zero; counters are int64.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import (BUCKET, DIRECT, I64, LDS, Instance, check, config, fill_splitmix, gen_index, lib)

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
_libm = C.CDLL("libm.so.6")
_libm.exp2.restype = C.c_double
_libm.exp2.argtypes = [C.c_double]


def _mix(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _at(seed: int, i: int) -> int:
    return _mix((seed + (i + 1) * GOLDEN) & M64)


def _u01(d: int) -> float:
    return float(d >> 11) * (2.0 ** -53)


@dataclass
class MixParams:
    kind: int
    dist: int
    table_len: int
    nops: int
    s: float
    seed: int


def mix_params(count: int, seed: int, lo_tab=10, hi_tab=24, lo_ops=12, hi_ops=22) -> list[MixParams]:
    out = []
    for t in range(count):
        sub = _at(seed, t)
        d = [_at(sub, k) for k in range(6)]
        tl = int(_libm.exp2(float(lo_tab) + float(hi_tab - lo_tab) * _u01(d[1])))
        no = int(_libm.exp2(float(lo_ops) + float(hi_ops - lo_ops) * _u01(d[2])))
        dist = (d[3] * 3) >> 64
        out.append(MixParams(kind=d[0] >> 63, dist=dist, table_len=tl, nops=no,
                             s=0.8 + 0.6 * _u01(d[4]), seed=d[5]))
    return out


class Mix:
    """Device arenas and the bale_instance array of one mix."""

    def __init__(self, params: list[MixParams], device, stream=None):
        import torch
        self.params = params
        tab = sum(p.table_len for p in params)
        ops = sum(p.nops for p in params)
        gops = sum(p.nops for p in params if p.kind == 1)
        self.tables = torch.empty(tab, dtype=torch.int64, device=device)
        self.index = torch.empty(ops, dtype=torch.int64, device=device)
        self.tgt = torch.empty(max(gops, 1), dtype=torch.int64, device=device)
        self.inst = (Instance * len(params))()
        self.views = []
        to = io = go = 0
        for i, p in enumerate(params):
            tv = self.tables[to:to + p.table_len]
            iv = self.index[io:io + p.nops]
            gv = self.tgt[go:go + p.nops] if p.kind == 1 else None
            if p.kind == 1:
                fill_splitmix(tv, stream)
            else:
                tv.zero_()
            gen_index(iv, p.table_len, dist=1 if p.dist == 1 else 0, s=p.s,
                      run_mean=64 if p.dist == 2 else 0, seed=p.seed, stream=stream)
            I = self.inst[i]
            I.kind, I.dtype, I.n, I.table_len = p.kind, I64, p.nops, p.table_len
            I.index, I.table = iv.data_ptr(), tv.data_ptr()
            I.tgt = gv.data_ptr() if gv is not None else None
            I.cfg = config()
            self.views.append((tv, iv, gv))
            to += p.table_len
            io += p.nops
            go += p.nops if p.kind == 1 else 0
        self.total_ops = ops

    def reset_counts(self) -> None:
        for p, (tv, _, _) in zip(self.params, self.views):
            if p.kind == 0:
                tv.zero_()

    def set_configs(self, cfgs) -> None:
        for I, c in zip(self.inst, cfgs):
            I.cfg = c

    def configs(self):
        return [BaleConfigCopy(I.cfg) for I in self.inst]


def BaleConfigCopy(c):
    return config(c.strategy, c.block, c.unroll, c.tile_log2, c.merge, c.groups_per_cu, c.hot,
                  c.task_log2)


class Batch:
    """A planned, graph-captured batch (bale_batch_create)."""

    def __init__(self, inst, count: int, stream=None):
        from . import _stream
        self.h = lib().bale_batch_create(inst, count, _stream(stream))
        if not self.h:
            from . import BaleError, last_error
            raise BaleError(-1, last_error())

    def launch(self, stream=None) -> None:
        from . import _stream
        check(lib().bale_batch_launch(self.h, _stream(stream)))

    def launch_direct(self, stream=None) -> None:
        from . import _stream
        check(lib().bale_batch_launch_direct(self.h, _stream(stream)))

    def stats(self) -> dict:
        out = (C.c_int64 * 6)()
        check(lib().bale_batch_stats(self.h, out, 6))
        keys = ["tasks_a", "tasks_scatter", "apply_slots", "sketches", "bucketed", "scratch_bytes"]
        return dict(zip(keys, list(out)))

    def close(self) -> None:
        if self.h:
            lib().bale_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_plain(inst, count: int, stream=None) -> None:
    from . import _stream
    check(lib().bale_batch_launch_plain(inst, count, _stream(stream)))


def launch_each(inst, count: int, stream=None) -> None:
    """Each instance's own configuration as one library call per instance."""
    from . import _stream
    check(lib().bale_batch_launch_each(inst, count, _stream(stream)))


def tune(inst, count: int, min_ops: int = 1 << 22, max_cand: int = 32, stream=None,
         chunk: int = 0, log=None):
    """Exhaustive autotuner: fills inst[i].cfg; returns the cost table
    (count x max_cand seconds per op, -1 where unused) as a list of lists.
    Instances are tuned independently; with `chunk` they go to the library
    `chunk` at a time and `log` (if given) hears after each."""
    from . import _stream
    costs = (C.c_float * (count * max_cand))()
    step = chunk if chunk > 0 else max(count, 1)
    for k0 in range(0, count, step):
        k = min(step, count - k0)
        part = C.cast(C.addressof(inst) + k0 * C.sizeof(Instance), C.POINTER(Instance))
        cpart = C.cast(C.addressof(costs) + k0 * max_cand * C.sizeof(C.c_float), C.POINTER(C.c_float))
        check(lib().bale_tune(part, k, min_ops, cpart, max_cand, _stream(stream)))
        if log is not None:
            log(f"tuned {k0 + k}/{count} instances")
    return [list(costs[i * max_cand:(i + 1) * max_cand]) for i in range(count)]


def candidates(kind: int, table_len: int):
    from . import BaleConfig
    buf = (BaleConfig * 64)()
    n = lib().bale_tune_candidates(kind, table_len, buf, 64)
    return [buf[i] for i in range(n)]


STRATEGIES = {DIRECT: "direct", LDS: "lds", BUCKET: "bucket"}
