"""Batched uneven instances: parameter generation (CPU) and the persistent,
graph-captured batch against the oracle (GPU)."""
import numpy as np
import pytest

import bale_old_amd as B
from bale_old_amd import mix as M


def test_mix_params_match_oracle(orc):
    got = M.mix_params(4096, 4)
    want = orc.batch_params(4096, 4)
    for g, w in zip(got, want):
        assert (g.kind, g.dist, g.table_len, g.nops, g.seed) == \
               (w["kind"], w["dist"], w["table_len"], w["nops"], w["seed"])
        assert g.s == w["s"]
    small = M.mix_params(50, 9, 8, 16, 10, 16)
    assert all(256 <= p.table_len < 65536 and 1024 <= p.nops < 65536 for p in small)


def test_candidates_cover_the_strategies():
    c_small = M.candidates(0, 1 << 12)
    strats = {c.strategy for c in c_small}
    assert strats == {B.DIRECT, B.LDS, B.BUCKET}
    assert {c.strategy for c in M.candidates(0, 1 << 20)} == {B.DIRECT, B.BUCKET}
    assert {c.strategy for c in M.candidates(1, 1 << 20)} == {B.DIRECT}
    # gathers from tables of at most 4096 entries can stage the table in LDS
    assert {c.strategy for c in M.candidates(1, 4096)} == {B.DIRECT, B.LDS}
    assert {c.strategy for c in M.candidates(1, 4097)} == {B.DIRECT}
    assert any(c.hot for c in c_small) and any(c.merge for c in c_small)


def _check_mix(mix, orc, times=1, only=None):
    for i, (p, (tv, iv, gv)) in enumerate(zip(mix.params, mix.views)):
        if only is not None and i not in only:
            continue
        idx = iv.cpu().numpy()
        if p.kind == 1:
            want = orc.ig_agi(idx, orc.fill_splitmix(p.table_len))
            assert np.array_equal(gv.cpu().numpy(), want), (i, p)
        else:
            want = np.zeros(p.table_len, np.int64)
            orc.histo_agi(idx, want)
            assert np.array_equal(tv.cpu().numpy(), times * want), (i, p, mix.inst[i].cfg)


@pytest.mark.gpu
def test_batch_every_candidate_matches_oracle(dev, orc):
    params = M.mix_params(96, 11, 8, 17, 10, 17)
    mix = M.Mix(params, dev)
    # give each instance a different candidate configuration, cycling through the list
    for i, p in enumerate(params):
        cands = M.candidates(p.kind, p.table_len)
        mix.inst[i].cfg = cands[i % len(cands)]
    b = M.Batch(mix.inst, len(params))
    st = b.stats()
    assert st["tasks_a"] > 0
    b.launch()
    B.sync_errors()
    _check_mix(mix, orc, 1)
    # replaying the graph adds again (histograms) and rewrites the same gathers
    b.launch()
    b.launch_direct()
    B.sync_errors()
    _check_mix(mix, orc, 3)
    b.close()


@pytest.mark.gpu
def test_batch_gathers_from_lds(dev, orc):
    """Gathers configured LDS read an LDS copy of tables of at most 4096
    entries (1, 7, 4096); a larger table (4097) runs the plain gather."""
    params = [M.MixParams(1, d, tl, n, 1.1, 9000 + k)
              for k, (tl, n, d) in enumerate([(1, 5000, 0), (7, 70000, 1), (4096, 1 << 18, 0),
                                               (4096, 333, 2), (4097, 1 << 17, 0)])]
    mix = M.Mix(params, dev)
    for i in range(len(params)):
        mix.inst[i].cfg = B.config(B.LDS, task_log2=14 + i % 3)
    b = M.Batch(mix.inst, len(params))
    b.launch()
    B.sync_errors()
    _check_mix(mix, orc, 1)
    b.close()


@pytest.mark.gpu
def test_batch_bucketed_and_hot_paths(dev, orc):
    params = M.mix_params(24, 5, 14, 20, 14, 19)
    mix = M.Mix(params, dev)
    for i, p in enumerate(params):
        mix.inst[i].cfg = B.config(B.BUCKET, hot=i % 2, merge=(i // 2) % 2, task_log2=12 + i % 4) \
            if p.kind == 0 else B.config(B.DIRECT, task_log2=14)
    b = M.Batch(mix.inst, len(params))
    assert b.stats()["bucketed"] == sum(1 for p in params if p.kind == 0)
    b.launch()
    B.sync_errors()
    _check_mix(mix, orc, 1)
    b.close()


@pytest.mark.gpu
def test_plain_and_tuned_batches(dev, orc):
    params = M.mix_params(40, 12, 8, 18, 10, 17)
    mix = M.Mix(params, dev)
    M.launch_plain(mix.inst, len(params))
    B.sync_errors()
    _check_mix(mix, orc, 1)
    costs = M.tune(mix.inst, len(params), min_ops=1 << 18)
    assert all(any(c > 0 for c in row) for row in costs)
    # tuning ran in scratch: the caller's counts are unchanged
    _check_mix(mix, orc, 1)
    b = M.Batch(mix.inst, len(params))
    b.launch()
    B.sync_errors()
    _check_mix(mix, orc, 2)
    b.close()


@pytest.mark.gpu
def test_batch_rejects_bad_instances(dev):
    import ctypes as C
    inst = (B.Instance * 1)()
    inst[0].kind = 3
    assert not B.lib().bale_batch_create(inst, 1, None)
    assert "instance 0" in B.last_error()
    assert B.lib().bale_batch_launch_plain(inst, 1, None) == B.ERR_ARG
