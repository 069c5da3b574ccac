"""Train the learned selector from the exhaustive autotuner's own timings.

    python -m bale_old_amd.train_selector --instances 20000 --out bale_old_amd/selector_weights.bin

or in parts (each part a whole number of chunks; the parts' seeds follow on):

    python -m bale_old_amd.train_selector --instances 10240 --collect-only --dump a.pt
    python -m bale_old_amd.train_selector --instances 9760 --seed0 1020 --collect-only --dump b.pt
    python -m bale_old_amd.train_selector --from-dump a.pt b.pt

Training instances come from mixes with seeds 1000, 1001, ... (disjoint from
the benchmark mixes, seeds 4 and 5), generated and tuned in chunks so that at
most one chunk's tables are resident.  For each instance the sketch features
are computed on the GPU and every candidate's timing is kept.  The default
loss (--loss cost) is the cross-entropy of the choice over the instance's
timed candidates against targets weighted by (best / time)^8; --loss ce trains
on the fastest candidate's classes instead.  --from-dump retrains from a
previous run's saved timings.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import torch

from . import lib, sync_errors
from . import mix as M
from . import selector as S


def collect(n_inst: int, chunk: int, device, min_ops: int, seed0: int = 1000, log=print):
    feats, labels, costs_all, kinds, tlens, nops = [], [], [], [], [], []
    done, seed = 0, seed0
    t0 = time.time()
    while done < n_inst:
        k = min(chunk, n_inst - done)
        params = M.mix_params(k, seed)
        mix = M.Mix(params, device)
        f = S.features(mix.inst, k, device)
        costs = M.tune(mix.inst, k, min_ops=min_ops)
        sync_errors()
        for i in range(k):
            labels.append(S.labels_of(mix.inst[i].cfg))
            kinds.append(params[i].kind)
            tlens.append(params[i].table_len)
            nops.append(params[i].nops)
        feats.append(f.cpu())
        costs_all.extend(costs)
        done += k
        seed += 1
        del mix
        torch.cuda.empty_cache()
        log(f"[train] collected {done}/{n_inst} instances ({time.time() - t0:.0f}s)")
    return (torch.cat(feats), torch.tensor(labels, dtype=torch.int32), costs_all, kinds, tlens,
            nops)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=20000)
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--min-ops", type=int, default=1 << 19)
    ap.add_argument("--seed0", type=int, default=1000,
                    help="first mix seed (collect in parts: a part of k chunks ends before seed0 + k)")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=str(S.WEIGHTS))
    ap.add_argument("--dump", default=None, help="also save features/labels/costs (torch.save)")
    ap.add_argument("--from-dump", nargs="+", default=None,
                    help="train from --dump files instead of collecting (parts in seed order, "
                         "together the seeds from 1000 in chunks of --chunk)")
    ap.add_argument("--collect-only", action="store_true", help="collect and --dump, do not train")
    ap.add_argument("--loss", choices=("cost", "ce"), default="cost")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    lib().bale_hip_init(0)
    if a.from_dump:
        parts = [torch.load(f, weights_only=True) for f in a.from_dump]
        feats = torch.cat([d["features"] for d in parts])
        labels = torch.cat([d["labels"] for d in parts])
        costs = [c for d in parts for c in d["costs"]]
        kinds = [k for d in parts for k in d["kinds"]]
        params, seed = [], 1000
        while len(params) < feats.shape[0]:
            params += M.mix_params(min(a.chunk, feats.shape[0] - len(params)), seed)
            seed += 1
        if [p.kind for p in params] != list(kinds):
            raise SystemExit(f"{a.from_dump}: instances do not match seeds 1000.. in chunks of {a.chunk}")
        tlens = [p.table_len for p in params]
        nops = [p.nops for p in params]
    else:
        feats, labels, costs, kinds, tlens, nops = collect(a.instances, a.chunk, dev, a.min_ops,
                                                           seed0=a.seed0, log=log)
        if a.collect_only:
            torch.save({"features": feats, "labels": labels, "costs": costs, "kinds": kinds,
                        "tlens": tlens, "nops": nops}, a.dump)
            log(f"[train] saved {feats.shape[0]} instances to {a.dump}")
            return 0
    n = feats.shape[0]
    n_val = max(1, n // 10)
    tr_f, tr_l = feats[n_val:].to(dev), labels[n_val:].to(dev)
    va_f, va_l = feats[:n_val].to(dev), labels[:n_val]
    if a.loss == "cost":
        cands, maxc = S.cand_table(kinds[n_val:], tlens[n_val:], costs[n_val:])
        P, loss = S.train_cost(tr_f, cands, maxc, steps=a.steps, batch=a.batch, lr=a.lr)
    else:
        P, loss = S.train(tr_f, tr_l, steps=a.steps, batch=a.batch, lr=a.lr)
    _, cls = S.infer(P, va_f)
    acc = (cls.cpu() == va_l).float().mean(dim=0).tolist()
    # regret on the validation split: sum of chosen candidates' costs over the
    # sum of the best costs (a ratio of total time, as the batch pays it)
    logits, _ = S.infer(P, va_f)
    rows = logits.cpu().tolist()
    chosen_t = best_t = 0.0
    for i in range(n_val):
        cands = M.candidates(kinds[i], tlens[i])
        c = S.choose(rows[i], cands)
        k = next(j for j, x in enumerate(cands) if S.labels_of(x) == S.labels_of(c))
        row = [x for x in costs[i][:len(cands)]]
        valid = [x for x in row if x > 0]
        if not valid or row[k] <= 0:
            continue
        chosen_t += row[k] * nops[i]
        best_t += min(valid) * nops[i]
    meta = {"instances": n, "validation": n_val, "steps": a.steps, "batch": a.batch, "lr": a.lr,
            "loss": a.loss,
            "final_loss": float(loss[-50:].mean()), "val_head_accuracy": acc,
            "val_time_vs_exhaustive": chosen_t / best_t if best_t else None,
            "label_counts": {h: torch.bincount(labels[:, h].long(), minlength=4).tolist()
                             for h in range(3)}}
    S.save(P, a.out, meta)
    if a.dump:
        torch.save({"features": feats, "labels": labels, "costs": costs, "kinds": kinds,
                    "tlens": tlens, "nops": nops}, a.dump)
    print(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
