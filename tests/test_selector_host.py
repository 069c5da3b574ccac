"""Host-side selector plumbing (no GPU): the candidate table the cost-aware
training reads."""
import math

import torch

from bale_old_amd import mix as M
from bale_old_amd import selector as S


def test_cand_table_relative_times():
    kinds = [0, 1, 0]
    tlens = [1 << 12, 1 << 20, 1 << 16]
    costs = []
    for k, tl in zip(kinds, tlens):
        n = len(M.candidates(k, tl))
        costs.append([2.0 + 0.5 * j if j % 3 else -1.0 for j in range(n)] + [9.0, 9.0])
    T, maxc = S.cand_table(kinds, tlens, costs)
    assert T.shape == (3, maxc, 4)
    for i, (k, tl) in enumerate(zip(kinds, tlens)):
        cands = M.candidates(k, tl)
        valid = [(S.labels_of(c), x) for c, x in zip(cands, costs[i]) if x > 0]
        best = min(x for _, x in valid)
        assert len(valid) <= maxc
        for j, (lab, x) in enumerate(valid):
            assert tuple(int(v) for v in T[i, j, :3].tolist()) == lab
            assert math.isclose(float(T[i, j, 3]), x / best, rel_tol=1e-6)
        assert bool((T[i, len(valid):, 0] == -1).all())
        assert float(T[i, :len(valid), 3].min()) == 1.0
