"""The learned selector's kernels: sketch features against a numpy
restatement of the same sample, MLP inference against a float32 numpy forward
pass (tolerance 1e-4 relative), and training on a learnable synthetic task."""
import math

import numpy as np
import pytest

import bale_old_amd as B
from bale_old_amd import mix as M
from bale_old_amd import selector as S

pytestmark = pytest.mark.gpu


def np_features(idx: np.ndarray, table_len: int, kind: int, dtype: int) -> np.ndarray:
    n = idx.size
    if n <= 4096:
        samp = idx.astype(np.uint64)
        nwin = 1
        win_start = [0]
    else:
        starts = [(w * (n - 64)) // 63 for w in range(64)]
        samp = np.concatenate([idx[s:s + 64] for s in starts]).astype(np.uint64)
        nwin = 64
    s32 = (samp & 0xFFFFFFFF).astype(np.uint32)
    m = s32.size
    eq = sum(1 for k in range(m) if k % 64 != 0 and s32[k] == s32[k - 1])
    _, counts = np.unique(s32, return_counts=True)
    p = counts / m
    f = np.zeros(12, np.float32)
    f[0] = math.log2(table_len) / 24
    f[1] = math.log2(n) / 22
    f[2] = kind
    f[3] = dtype
    f[4] = min(2, max(-2, (math.log2(n) - math.log2(table_len)) / 12))
    f[5] = float(-(p * np.log2(p)).sum()) / 12
    f[6] = np.sort(counts)[::-1][:16].sum() / m
    f[7] = counts.max() / m
    f[8] = counts.size / m
    f[9] = eq / m
    runs = max(1.0, nwin + (m - nwin - eq))
    f[10] = math.log2(m / runs) / 6
    f[11] = counts[counts >= 2].sum() / m
    return f


def test_sketch_features(dev):
    params = [M.MixParams(0, 0, 1 << 20, 1 << 18, 1.0, 1), M.MixParams(0, 1, 1 << 16, 1 << 17, 1.2, 2),
              M.MixParams(1, 2, 1 << 14, 100_000, 1.0, 3), M.MixParams(0, 0, 1000, 3000, 1.0, 4),
              M.MixParams(1, 1, 1 << 12, 5000, 1.4, 5)]
    mix = M.Mix(params, dev)
    f = S.features(mix.inst, len(params), dev).cpu().numpy()
    for i, p in enumerate(params):
        want = np_features(mix.views[i][1].cpu().numpy(), p.table_len, p.kind, 0)
        np.testing.assert_allclose(f[i], want, rtol=1e-4, atol=1e-5, err_msg=str(p))


def np_forward(P: np.ndarray, x: np.ndarray) -> np.ndarray:
    W1 = P[:768].reshape(64, 12); b1 = P[768:832]
    W2 = P[832:832 + 4096].reshape(64, 64); b2 = P[4928:4992]
    W3 = P[4992:4992 + 640].reshape(10, 64); b3 = P[5632:5642]
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def test_mlp_inference_matches_numpy(dev):
    import torch
    assert S.param_count() == 5642
    P = S.init_params(3, dev)
    x = torch.rand((300, 12), device=dev)
    logits, cls = S.infer(P, x)
    want = np_forward(P.cpu().numpy().astype(np.float32), x.cpu().numpy().astype(np.float32))
    np.testing.assert_allclose(logits.cpu().numpy(), want, rtol=1e-4, atol=1e-4)
    for h, (off, n) in enumerate(S.HEADS):
        assert np.array_equal(cls[:, h].cpu().numpy(), want[:, off:off + n].argmax(axis=1))


def test_mlp_training_learns(dev):
    import torch
    g = torch.Generator().manual_seed(0)
    x = torch.rand((4000, 12), generator=g)
    lab = torch.stack([(x[:, 0] > 0.5).long() + (x[:, 1] > 0.7).long(),
                       (x[:, 2] * 4).long().clamp(max=3),
                       ((x[:, 3] + x[:, 4]) * 1.5).long().clamp(max=2)], dim=1).int()
    P, loss = S.train(x.to(dev), lab.to(dev), steps=1500, batch=256, lr=5e-3)
    assert loss[-100:].mean() < 0.5 * loss[:20].mean()
    _, cls = S.infer(P, x.to(dev))
    acc = (cls.cpu() == lab).float().mean(dim=0)
    assert (acc > 0.85).all(), acc
    # the choice step picks the top-scoring valid candidate
    logits, _ = S.infer(P, x[:5].to(dev))
    cands = M.candidates(0, 1 << 12)
    c = S.choose(logits[0].cpu().tolist(), cands)
    assert any(S.labels_of(c) == S.labels_of(y) for y in cands)


def test_weights_roundtrip(dev, tmp_path):
    P = S.init_params(5, dev)
    S.save(P, tmp_path / "w.bin", {"note": "test"})
    Q, meta = S.load(dev, tmp_path / "w.bin")
    assert meta == {"note": "test"}
    assert bool((P == Q).all())


def test_cost_training_folds_standardisation(dev):
    """steps=0: the returned parameters are the seeded initialisation with the
    feature standardisation folded into the first layer, so raw features
    through them equal standardised features through the initialisation."""
    import torch
    g = torch.Generator().manual_seed(1)
    x = torch.rand((500, 12), generator=g) * torch.linspace(0.5, 30, 12) + torch.linspace(-3, 3, 12)
    cands = torch.tensor([[[0, 0, 0, 1.0], [1, 2, 1, 1.5]]] * 500, dtype=torch.float32)
    P, _ = S.train_cost(x.to(dev), cands, 2, steps=0)
    P0 = S.init_params(7, dev).cpu().numpy().astype(np.float64)
    xs = x.numpy().astype(np.float64)
    z = (xs - xs.mean(0)) / (xs.std(0) + 1e-6)
    want = np_forward(P0, z)
    got = np_forward(P.cpu().numpy().astype(np.float64), xs)
    np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-3)


def test_cost_training_minimises_expected_time(dev):
    """Candidates whose relative time depends on the features: training on the
    expected relative time must pick near-best candidates, including where the
    best one differs from head to head."""
    import torch
    g = torch.Generator().manual_seed(2)
    n = 4000
    x = torch.rand((n, 12), generator=g)
    rows = []
    for i in range(n):
        a, b = float(x[i, 0]), float(x[i, 1])
        # (strategy, variant, task, relative time)
        r = [(0, 0, 1, 1.0 + 2 * a), (1, 1, 1, 1.0 + 2 * (1 - a)), (2, 2, 0, 1.0 + 3 * abs(b - 0.5)),
             (2, 3, 2, 1.2)]
        best = min(t for *_, t in r)
        rows.append([(s, v, t, tt / best) for s, v, t, tt in r])
    cands = torch.tensor(rows, dtype=torch.float32)
    P, loss = S.train_cost(x.to(dev), cands, 4, steps=1500, batch=256, lr=5e-3)
    assert loss[-100:].mean() < loss[:20].mean()
    logits, _ = S.infer(P, x.to(dev))
    lp = logits.cpu().double()
    tot_chosen = tot_best = 0.0
    for i in range(n):
        sc = []
        for s, v, t, r in rows[i]:
            z = [torch.log_softmax(lp[i, o:o + k], 0) for o, k in S.HEADS]
            sc.append(float(z[0][int(s)] + z[1][int(v)] + z[2][int(t)]))
        j = int(np.argmax(sc))
        tot_chosen += rows[i][j][3]
        tot_best += 1.0
    assert tot_chosen / tot_best < 1.05, tot_chosen / tot_best
