"""Learned strategy selector: sketch features -> fp32 MLP -> per-instance config.

The sketch, inference and training run on the GPU in libbale_hip.so
(bale_sketch, bale_mlp_infer, bale_mlp_train, bale_mlp_train_cost).  This module maps between
tuner candidates and the MLP's three heads, trains from the exhaustive
tuner's timings, and stores / loads the weights.

Heads: strategy {direct, lds, bucket}; variant {plain, runs, hot, hot+runs};
task size {2^14, 2^16, 2^18}.  A prediction is turned into a configuration by
scoring each of the instance's valid candidates with the sum of the three
heads' log-probabilities and taking the best.
"""
from __future__ import annotations

import ctypes as C
import math
import struct
from pathlib import Path

from . import BUCKET, DIRECT, LDS, BaleConfig, check, config, lib

N_FEAT = 12
N_OUT = 10
HEADS = ((0, 3), (3, 4), (7, 3))
STRAT_IDX = {DIRECT: 0, LDS: 1, BUCKET: 2}
TASKS = (14, 16, 18)
WEIGHTS = Path(__file__).resolve().parent / "selector_weights.bin"
MAGIC = b"BALEMLP1"


def _proto():
    L = lib()
    if getattr(L, "_sel_proto", False):
        return L
    VP, I64 = C.c_void_p, C.c_int64
    L.bale_mlp_param_count.restype = C.c_int
    L.bale_sketch.restype = C.c_int
    L.bale_sketch.argtypes = [VP, I64, VP, VP]
    L.bale_mlp_infer.restype = C.c_int
    L.bale_mlp_infer.argtypes = [VP, VP, I64, VP, VP, VP]
    L.bale_mlp_train.restype = C.c_int
    L.bale_mlp_train.argtypes = [VP, VP, VP, VP, I64, C.c_int, C.c_int, C.c_float, C.c_uint64, VP, VP]
    L.bale_mlp_train_cost.restype = C.c_int
    L.bale_mlp_train_cost.argtypes = [VP, VP, VP, VP, C.c_int, I64, C.c_int, C.c_int, C.c_float,
                                      C.c_uint64, VP, VP, VP]
    L._sel_proto = True
    return L


def param_count() -> int:
    return _proto().bale_mlp_param_count()


def labels_of(cfg) -> tuple[int, int, int]:
    s = STRAT_IDX.get(cfg.strategy, 0)
    v = (2 if cfg.hot else 0) + (1 if cfg.merge > 0 else 0)
    t = TASKS.index(cfg.task_log2) if cfg.task_log2 in TASKS else 1
    return s, v, t


def features(inst, count: int, device, stream=None):
    """count x 12 float32 tensor of sketch features (on the GPU)."""
    import torch
    from . import _stream
    f = torch.empty((count, N_FEAT), dtype=torch.float32, device=device)
    check(_proto().bale_sketch(C.addressof(inst), count, f.data_ptr(), _stream(stream)))
    return f


def init_params(seed: int, device):
    """He-style uniform initialisation, seeded, built on the host."""
    import torch
    g = torch.Generator().manual_seed(seed)
    P = torch.empty(param_count(), dtype=torch.float32)
    off = 0
    for fan_in, fan_out in ((N_FEAT, 64), (64, 64), (64, N_OUT)):
        bound = math.sqrt(6.0 / fan_in)
        P[off:off + fan_in * fan_out] = (torch.rand(fan_in * fan_out, generator=g) * 2 - 1) * bound
        off += fan_in * fan_out
        P[off:off + fan_out] = 0.0
        off += fan_out
    assert off == P.numel()
    return P.to(device)


def train(feats, labels, *, steps=3000, batch=512, lr=3e-3, seed=7, params=None, stream=None):
    """Train on the GPU; returns (params, per-step loss)."""
    import torch
    from . import _stream
    dev = feats.device
    P = init_params(seed, dev) if params is None else params.clone()
    adam = torch.zeros(2 * P.numel(), dtype=torch.float32, device=dev)
    loss = torch.zeros(max(steps, 1), dtype=torch.float32, device=dev)
    lab = labels.to(device=dev, dtype=torch.int32).contiguous()
    check(_proto().bale_mlp_train(P.data_ptr(), adam.data_ptr(), feats.contiguous().data_ptr(),
                                  lab.data_ptr(), feats.shape[0], steps, batch, lr, seed,
                                  loss.data_ptr(), _stream(stream)))
    torch.cuda.synchronize()
    return P, loss.cpu()


def cand_table(kinds, tlens, costs):
    """(count x maxc x 4 float tensor, maxc): for each instance the head classes
    of every candidate the tuner timed and its time over the instance's best;
    rows past an instance's candidates have strategy -1."""
    import torch
    from .mix import candidates
    cache = {}
    rows = []
    maxc = 1
    for kind, tl, cs in zip(kinds, tlens, costs):
        key = (kind, tl)
        if key not in cache:
            cache[key] = [labels_of(c) for c in candidates(kind, tl)]
        labs = cache[key]
        valid = [x for x in cs[:len(labs)] if x > 0]
        best = min(valid) if valid else 1.0
        row = [(a, b, c, x / best) for (a, b, c), x in zip(labs, cs) if x > 0]
        rows.append(row)
        maxc = max(maxc, len(row))
    T = torch.full((len(rows), maxc, 4), -1.0, dtype=torch.float32)
    for i, row in enumerate(rows):
        if row:
            T[i, :len(row)] = torch.tensor(row, dtype=torch.float32)
    return T, maxc


def train_cost(feats, cands, maxc, *, steps=4000, batch=512, lr=3e-3, seed=7, params=None,
               stream=None):
    """Train on the tuner's timings of every candidate (GPU): cross-entropy of
    the choice over an instance's candidates against targets weighted by
    (best / time)^8.  Features are standardised on the device and the
    standardisation folded into the first layer.  Returns (params, per-step
    loss)."""
    import torch
    from . import _stream
    dev = feats.device
    P = init_params(seed, dev) if params is None else params.clone()
    adam = torch.zeros(2 * P.numel(), dtype=torch.float32, device=dev)
    loss = torch.zeros(max(steps, 1), dtype=torch.float32, device=dev)
    n = feats.shape[0]
    scratch = torch.empty(2 * N_FEAT + n * N_FEAT, dtype=torch.float32, device=dev)
    cd = cands.to(device=dev, dtype=torch.float32).contiguous()
    check(_proto().bale_mlp_train_cost(P.data_ptr(), adam.data_ptr(), feats.contiguous().data_ptr(),
                                       cd.data_ptr(), maxc, n, steps, batch, lr, seed,
                                       loss.data_ptr(), scratch.data_ptr(), _stream(stream)))
    torch.cuda.synchronize()
    return P, loss.cpu()


def infer(P, feats, stream=None):
    """(logits count x 10, argmax classes count x 3), on the GPU."""
    import torch
    from . import _stream
    n = feats.shape[0]
    logits = torch.empty((n, N_OUT), dtype=torch.float32, device=feats.device)
    cls = torch.empty((n, 3), dtype=torch.int32, device=feats.device)
    check(_proto().bale_mlp_infer(P.data_ptr(), feats.contiguous().data_ptr(), n, logits.data_ptr(),
                                  cls.data_ptr(), _stream(stream)))
    return logits, cls


def choose(logits_row, cands) -> BaleConfig:
    """Best valid candidate under the heads' summed log-probabilities."""
    lp = []
    for off, n in HEADS:
        z = logits_row[off:off + n]
        m = max(z)
        s = math.log(sum(math.exp(x - m) for x in z)) + m
        lp.append([x - s for x in z])
    best, bs = None, -1e30
    for c in cands:
        s, v, t = labels_of(c)
        sc = lp[0][s] + lp[1][v] + lp[2][t]
        if sc > bs:
            best, bs = c, sc
    return best


def select(inst, count: int, P, device, stream=None):
    """Learned configuration for every instance (also written into inst[i].cfg)."""
    from .mix import candidates
    f = features(inst, count, device, stream)
    logits, _ = infer(P, f, stream)
    rows = logits.cpu().tolist()
    out = []
    cache = {}
    for i in range(count):
        key = (inst[i].kind, inst[i].table_len <= (1 << 15), inst[i].table_len >= (1 << 12))
        if key not in cache:
            cache[key] = candidates(inst[i].kind, inst[i].table_len)
        c = choose(rows[i], cache[key])
        inst[i].cfg = c
        out.append(c)
    return out


def save(P, path: Path = WEIGHTS, meta: dict | None = None) -> None:
    import json
    data = P.detach().float().cpu().numpy().astype("<f4").tobytes()
    m = json.dumps(meta or {}).encode()
    with open(path, "wb") as fh:
        fh.write(MAGIC + struct.pack("<ii", P.numel(), len(m)) + m + data)


def load(device, path: Path = WEIGHTS):
    import json
    import numpy as np
    import torch
    raw = Path(path).read_bytes()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not a selector weights file")
    n, mlen = struct.unpack("<ii", raw[8:16])
    if n != param_count():
        raise ValueError(f"{path}: {n} parameters, the library expects {param_count()}")
    meta = json.loads(raw[16:16 + mlen].decode())
    arr = np.frombuffer(raw[16 + mlen:16 + mlen + 4 * n], dtype="<f4").copy()
    return torch.from_numpy(arr).to(device), meta


def default_config() -> BaleConfig:
    return config()
