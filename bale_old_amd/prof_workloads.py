"""One dispatch of each measured kernel, for hardware-counter runs:

    rocprofv3 --pmc FETCH_SIZE -- python3 bale_old_amd/prof_workloads.py gather copy histo

  gather  k_gather_direct over 2^28 uniform requests into a 2^28-entry table
  copy    k_probe_copy of exactly 1 GiB (read 1 GiB, write 1 GiB): the
          calibration that converts the counters to bytes
  histo   the planned (bucketed) histogram of 2^28 uniform updates into 2^26 int64 bins
  zipf    the same plan on 2^28 Zipf(1.1) updates (its sample switches the hot table on)
  small   the planned histogram of 2^24 uniform updates into 2^20 int64 bins
          (8 MiB of counters, small enough to stay in the caches)
  batch   the held-out mix of 4096 instances (seed 5), strategies from the learned
          selector, captured once and launched three times
"""
from __future__ import annotations

import sys
from pathlib import Path

import torch

if __package__ in (None, ""):  # run by file path: make the package importable
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
import bale_old_amd as B  # noqa: E402


def main(argv=None) -> int:
    which = set(argv if argv is not None else sys.argv[1:]) or {"gather", "copy", "histo"}
    dev = torch.device("cuda:0")
    B.check(B.lib().bale_hip_init(0))
    st = torch.cuda.current_stream()
    n = 1 << 28
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    if which & {"gather", "copy"}:
        table = torch.empty(n, dtype=torch.int64, device=dev)
        B.fill_splitmix(table)
        tgt = torch.empty(n, dtype=torch.int64, device=dev)
        if "gather" in which:
            B.gen_index(idx, n, seed=2)
            B.gather(tgt, idx, table, cfg=B.gather_plan(n, n))
        if "copy" in which:
            B.check(B.lib().bale_probe_copy(tgt.data_ptr(), table.data_ptr(), 1 << 30, st.cuda_stream))
        torch.cuda.synchronize()
        del table, tgt
    if "histo" in which:
        counts = torch.zeros(1 << 26, dtype=torch.int64, device=dev)
        B.gen_index(idx, 1 << 26, seed=3)
        B.histo(idx, counts, cfg=B.histo_plan(n, 1 << 26))
        del counts
    if "zipf" in which:
        counts = torch.zeros(1 << 26, dtype=torch.int64, device=dev)
        B.gen_index(idx, 1 << 26, dist=1, s=1.1, seed=3)
        B.histo(idx, counts, cfg=B.histo_plan(n, 1 << 26))
        del counts
    if "small" in which:
        m, nb = 1 << 24, 1 << 20
        counts = torch.zeros(nb, dtype=torch.int64, device=dev)
        B.gen_index(idx[:m], nb, seed=1)
        B.histo(idx[:m], counts, cfg=B.histo_plan(m, nb))
        del counts
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
    print("prof_workloads done:", " ".join(sorted(which)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
