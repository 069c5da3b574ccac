"""The five benchmark workloads at their full sizes, checked exactly.

Histograms and gathers at 2^24..2^28 elements are compared element for element
with independent computations of the same result: the oracle's generator and
histogram (2^24 updates), numpy.bincount of the device-generated indices on
the host (2^28 updates), and splitmix64 restated in torch on the GPU for the
2^28-request gather, whose table is table[j] = splitmix64(j).  Invariants
(sum of counts, int32 == int64, accumulation across strategies) add
size-independent checks.  The batched mixes run all 4096 instances.
"""
import numpy as np
import pytest
import torch

import bale_old_amd as B
import torch_check as TC


# ---- the torch restatement of splitmix64 (CPU, pinned) --------------------------
def test_torch_splitmix_matches_oracle_and_published(orc):
    j = torch.tensor([0, 1, 2, 3, (1 << 28) - 1, (1 << 62) + 12345, -1, -(1 << 63)],
                     dtype=torch.int64)
    got = [int(v) & TC.M64 for v in TC.splitmix64(j)]
    want = [orc.splitmix64(int(v) & TC.M64) for v in j]
    assert got == want
    # published outputs of the stream seeded with 0: the state after k steps
    # is k * golden, so splitmix64(k * golden - golden) is output k (0-based)
    pub = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    states = torch.tensor([TC._s64(k * 0x9E3779B97F4A7C15) for k in range(3)], dtype=torch.int64)
    assert [int(v) & TC.M64 for v in TC.splitmix64(states)] == pub
    r = torch.from_numpy(orc.gen_uniform(1 << 12, 1 << 40, 9))
    want = np.array([orc.splitmix64(int(x)) for x in r.tolist()], dtype=np.uint64).view(np.int64)
    assert np.array_equal(TC.splitmix64(r).numpy(), want)
    # a dense range against the oracle's table fill
    assert np.array_equal(TC.splitmix64(torch.arange(1 << 16, dtype=torch.int64)).numpy(),
                          orc.fill_splitmix(1 << 16))


def test_torch_gather_check_catches_a_wrong_element():
    idx = torch.arange(1000, dtype=torch.int64)
    tgt = TC.splitmix64(idx)
    assert TC.gather_matches_splitmix(tgt, idx, chunk=128)
    tgt[777] ^= 1 << 40
    assert not TC.gather_matches_splitmix(tgt, idx, chunk=128)


# ---- config: plain histogram, 2^24 updates into 2^20 int64 counters -------------
@pytest.mark.gpu
def test_plain_histogram_full(dev, orc):
    n, nbins = 1 << 24, 1 << 20
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(idx, nbins, seed=1)
    idx_h = idx.cpu().numpy()
    assert np.array_equal(idx_h, orc.gen_index(n, nbins, 0, 1.0, 0, 1))
    want = np.zeros(nbins, np.int64)
    orc.histo_agi(idx_h, want)
    plans = [B.config(B.DIRECT), B.config(B.DIRECT, merge=2), B.config(B.LDS, block=1024),
             B.config(B.BUCKET, block=1024), B.config(B.BUCKET, block=1024, hot=2), None]
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    for k, cfg in enumerate(plans):
        assert B.histo(idx, counts, cfg=cfg) == 0
        assert np.array_equal(counts.cpu().numpy(), (k + 1) * want), cfg
    del idx, counts
    torch.cuda.empty_cache()


# ---- config: plain index gather, 2^28 requests into a 2^28-entry table ----------
@pytest.mark.gpu
def test_plain_gather_full(dev, orc):
    n = tlen = 1 << 28
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(idx, tlen, seed=2)
    # the generator, element for element over all 2^28 requests
    assert np.array_equal(idx.cpu().numpy(), orc.gen_index(n, tlen, 0, 1.0, 0, 2))
    # the table against the torch restatement
    assert TC.gather_matches_splitmix(table, torch.arange(tlen, dtype=torch.int64, device=dev))
    tgt = torch.empty(n, dtype=torch.int64, device=dev)
    for cfg in (None, B.config(B.DIRECT, block=1024, unroll=8), B.config(B.BUCKET, block=1024),
                B.config(B.BUCKET, tile_log2=17, task_log2=10, groups_per_cu=2)):
        tgt.fill_(-7)
        B.gather(tgt, idx, table, cfg=cfg)
        assert TC.gather_matches_splitmix(tgt, idx), cfg
    # packed indices (the buffered models' input at one PE)
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    for cfg in (None, B.config(B.BUCKET)):
        tgt.fill_(-7)
        B.gather(tgt, pck, table, packed=True, cfg=cfg)
        assert TC.gather_matches_splitmix(tgt, idx), cfg
    del table, idx, tgt, pck
    torch.cuda.empty_cache()


# ---- beyond 2^30: the bucketed passes run in slices ---------------------------
@pytest.mark.gpu
def test_calls_past_2p30_run_in_slices(dev):
    """A gather and a histogram of 2^30 + 12345 requests / updates: the
    bucketed passes keep 32-bit positions and run in slices of 2^30, so the
    boundary falls inside both calls.  Checked element for element (splitmix64
    in torch; torch.bincount)."""
    n, tlen = (1 << 30) + 12345, 1 << 26
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(idx, tlen, seed=11)
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    tgt = torch.empty(n, dtype=torch.int64, device=dev)
    tgt.fill_(-7)
    B.gather(tgt, idx, table, cfg=B.config(B.BUCKET))
    assert B.lib().bale_hip_last_strategy() == B.BUCKET
    assert TC.gather_matches_splitmix(tgt, idx)
    del tgt, table
    torch.cuda.empty_cache()
    counts = torch.zeros(tlen, dtype=torch.int64, device=dev)
    B.histo(idx, counts, cfg=B.config(B.BUCKET, block=1024, hot=2))
    assert B.lib().bale_hip_last_strategy() == B.BUCKET
    assert torch.equal(counts, torch.bincount(idx, minlength=tlen))
    del idx, counts
    torch.cuda.empty_cache()


# ---- config: skewed histogram, 2^28 Zipf(1.1) updates into 2^26 bins -----------
@pytest.mark.gpu
@pytest.mark.parametrize("run_mean", [0, 64], ids=["zipf1.1", "zipf1.1_runs64"])
def test_skewed_histogram_full(dev, orc, run_mean):
    n, nbins = 1 << 28, 1 << 26
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(idx, nbins, dist=1, s=1.1, run_mean=run_mean, seed=3)
    idx_h = idx.cpu().numpy()
    head = 1 << 22
    assert np.array_equal(idx_h[:head], orc.gen_index(head, nbins, 1, 1.1, run_mean, 3))
    want = np.bincount(idx_h, minlength=nbins)
    del idx_h
    assert want.size == nbins and int(want.sum()) == n
    assert want.max() < (1 << 31)  # int32 counters cannot overflow on this input
    plans = [None, B.config(B.BUCKET, block=1024, hot=0), B.config(B.BUCKET, block=1024, hot=1, merge=1),
             B.config(B.DIRECT, merge=2, hot=1)]
    for dtype in (torch.int64, torch.int32):
        counts = torch.zeros(nbins, dtype=dtype, device=dev)
        for k, cfg in enumerate(plans):
            assert B.histo(idx, counts, cfg=cfg) == 0
            got = counts.cpu().numpy().astype(np.int64)
            assert np.array_equal(got, (k + 1) * want), (dtype, cfg)
        del counts
        torch.cuda.empty_cache()
    del idx
    torch.cuda.empty_cache()


# ---- configs: batched uneven mixes, 4096 instances -------------------------------
def _check_mix_full(mix, times):
    """Every instance: gathers against torch splitmix64, histograms against
    torch.bincount of the instance's indices (times launches)."""
    for i, (p, (tv, iv, gv)) in enumerate(zip(mix.params, mix.views)):
        if p.kind == 1:
            assert TC.gather_matches_splitmix(gv, iv), (i, p)
        else:
            want = torch.bincount(iv, minlength=p.table_len) * times
            assert torch.equal(tv, want), (i, p, mix.inst[i].cfg)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [4, 5])
def test_batched_mix_full(dev, orc, seed):
    from bale_old_amd import mix as MX
    from bale_old_amd import selector as S
    params = MX.mix_params(4096, seed)
    want = orc.batch_params(4096, seed)
    assert [(p.kind, p.table_len, p.nops) for p in params] == \
           [(w["kind"], w["table_len"], w["nops"]) for w in want]
    mix = MX.Mix(params, dev)
    # the instance generators: a few instances element for element
    for i in (0, 1, 4095):
        p = params[i]
        got = mix.views[i][1].cpu().numpy()
        ref = orc.gen_index(p.nops, p.table_len, 1 if p.dist == 1 else 0, p.s,
                            64 if p.dist == 2 else 0, p.seed)
        assert np.array_equal(got, ref), i
    # plain per-instance launches
    MX.launch_plain(mix.inst, 4096)
    B.sync_errors()
    _check_mix_full(mix, 1)
    # one graph launch with the default rule, then with the learned selector
    b = MX.Batch(mix.inst, 4096)
    b.launch()
    B.sync_errors()
    b.close()
    _check_mix_full(mix, 2)
    P, _ = S.load(dev)
    S.select(mix.inst, 4096, P, dev)
    strategies = {c.strategy for c in mix.configs()}
    assert len(strategies) >= 2  # the selector mixes strategies on this mix
    b = MX.Batch(mix.inst, 4096)
    b.launch()
    b.launch()
    B.sync_errors()
    b.close()
    _check_mix_full(mix, 4)
    del mix
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_batched_mix_tuned_full(dev):
    """The seed-4 mix with the exhaustive tuner's own choices: every 8th
    instance tuned (the whole mix takes minutes; the bench tunes all of it and
    checks every instance the same way), the rest on the default rule, one
    graph launched twice, every instance checked."""
    from bale_old_amd import mix as MX
    params = MX.mix_params(4096, 4)
    mix = MX.Mix(params, dev)
    sub = list(range(0, 4096, 8))
    sub_inst = (MX.Instance * len(sub))(*[mix.inst[i] for i in sub])
    MX.tune(sub_inst, len(sub), min_ops=1 << 17)
    for k, i in enumerate(sub):
        mix.inst[i].cfg = sub_inst[k].cfg
    kinds = {(c.strategy, c.hot, c.merge > 0) for i, c in enumerate(mix.configs()) if i % 8 == 0}
    assert len(kinds) >= 3  # the tuner picks several strategies on this mix
    b = MX.Batch(mix.inst, 4096)
    b.launch()
    b.launch()
    B.sync_errors()
    b.close()
    _check_mix_full(mix, 2)
    del mix
    torch.cuda.empty_cache()
