"""GPU parity: libbale_hip.so against the CPU oracle on the same seeded inputs.

Integer results must be bit-exact.  Inputs are generated on the GPU by the
library's generator, which is itself checked element for element against the
oracle's generator first.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import bale_old_amd as B

pytestmark = pytest.mark.gpu


def to_np(t):
    return t.cpu().numpy()


def gen(dev, n, nbins, dist=0, s=1.0, run_mean=0, seed=1):
    out = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(out, nbins, dist=dist, s=s, run_mean=run_mean, seed=seed)
    return out


# ---- generators ------------------------------------------------------------------
@pytest.mark.parametrize("nbins,dist,s,run_mean,seed", [
    (1 << 20, 0, 1.0, 0, 1),
    (1000003, 0, 1.0, 0, 2),
    (1 << 26, 1, 1.1, 0, 3),
    (1 << 26, 1, 1.1, 64, 3),
    (5000, 1, 0.8, 0, 4),
    (77777, 1, 1.4, 0, 5),
    (1 << 20, 0, 1.0, 64, 6),
    (123457, 1, 1.0, 16, 7),
])
def test_generator_matches_oracle(dev, orc, nbins, dist, s, run_mean, seed):
    n = 1 << 18
    got = to_np(gen(dev, n, nbins, dist, s, run_mean, seed))
    want = orc.gen_index(n, nbins, dist, s, run_mean, seed)
    assert np.array_equal(got, want)


def test_fill_and_pack_match_oracle(dev, orc):
    t = torch.empty(100_003, dtype=torch.int64, device=dev)
    B.fill_splitmix(t)
    assert np.array_equal(to_np(t), orc.fill_splitmix(100_003))
    B.fill_neg(t)
    assert np.array_equal(to_np(t), orc.fill_neg(100_003))
    idx = gen(dev, 5000, 1 << 30)
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    assert np.array_equal(to_np(pck), orc.pack(to_np(idx), 1))


# ---- histogram -------------------------------------------------------------------
HISTO_CASES = [
    # (nbins, n, dist, s, run_mean)
    (1, 1000, 0, 1.0, 0),
    (7, 12345, 0, 1.0, 0),
    (1000, 1 << 18, 1, 1.1, 0),
    (1 << 15, 1 << 19, 0, 1.0, 0),
    ((1 << 15) + 3, 1 << 19, 1, 1.3, 0),
    (1 << 20, 1 << 20, 0, 1.0, 64),
    (3_000_017, 1 << 20, 1, 1.1, 0),
]
STRATS = [(B.DIRECT, 0, 0), (B.DIRECT, 1, 0), (B.DIRECT, 2, 0), (B.DIRECT, 0, 1),
          (B.LDS, 0, 0), (B.LDS, 1, 0), (B.LDS, 2, 0),
          (B.BUCKET, 0, 0), (B.BUCKET, 0, 1), (B.BUCKET, 1, 1), (B.BUCKET, 0, 2),
          (B.AUTO, 0, 0)]


@pytest.mark.parametrize("case", HISTO_CASES, ids=lambda c: f"b{c[0]}_n{c[1]}_d{c[2]}_r{c[4]}")
@pytest.mark.parametrize("dtype", [torch.int64, torch.int32])
def test_histo_strategies_match_oracle(dev, orc, case, dtype):
    nbins, n, dist, s, run_mean = case
    idx = gen(dev, n, nbins, dist, s, run_mean, seed=nbins + n)
    idx_h = to_np(idx)
    base = torch.randint(0, 1000, (nbins,), dtype=dtype, device=dev)
    want = to_np(base).astype(np.int64 if dtype == torch.int64 else np.int32)
    orc.histo_agi(idx_h, want)
    for strat, merge, hot in STRATS:
        for unroll, block in ((4, 256), (16, 1024)):
            if strat == B.BUCKET and unroll != 4:
                continue
            counts = base.clone()
            cfg = B.config(strat, block=block, unroll=unroll, merge=merge, hot=hot)
            assert B.histo(idx, counts, cfg=cfg) == 0
            assert np.array_equal(to_np(counts), want), (strat, merge, hot, unroll, block)


@pytest.mark.parametrize("tile_log2", [9, 12, 15, 16])
def test_histo_tiles(dev, orc, tile_log2):
    nbins, n = 100_003, 1 << 19
    idx = gen(dev, n, nbins, 1, 1.05, 8, seed=tile_log2)
    want = np.bincount(to_np(idx), minlength=nbins)
    for strat, hot in ((B.LDS, 0), (B.BUCKET, 0), (B.BUCKET, 1), (B.BUCKET, 2)):
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        B.histo(idx, counts, cfg=B.config(strat, block=1024, tile_log2=tile_log2, hot=hot))
        assert np.array_equal(to_np(counts), want), (strat, hot)


@pytest.mark.parametrize("case", ["one_bin", "word_pair", "zipf1.4", "runs64"])
@pytest.mark.parametrize("dtype", [torch.int64, torch.int32])
def test_histo_packed_tile_spills(dev, case, dtype):
    """2^16-bin buckets keep two 16-bit LDS counters per word; a counter that
    wraps is owed 2^16 and a low-half wrap's carry is taken off the high bin:
    drive single bins, and both bins of one word, far past 2^16 adds within one
    bucket segment, and the last bin of an odd-width bucket."""
    if case == "one_bin":
        nbins, n = (1 << 17) + 3, 1 << 21
        idx = torch.full((n,), 70001, dtype=torch.int64, device=dev)
        idx[::3] = 131074  # second bucket, last (partial) tile
    elif case == "word_pair":
        # bins 70000 / 70001 share a word, 131074 is the odd-width tile's
        # last bin (low half; its high half is no bin)
        nbins, n = (1 << 17) + 3, 3 << 20
        idx = torch.full((n,), 70000, dtype=torch.int64, device=dev)
        idx[1::3] = 70001
        idx[2::6] = 131074
        idx[5::12] = 131073
    elif case == "zipf1.4":
        nbins, n = (1 << 18) + 5, 1 << 21
        idx = gen(dev, n, nbins, 1, 1.4, 0, seed=17)
    else:
        nbins, n = (1 << 18) + 5, 1 << 21
        idx = gen(dev, n, nbins, 1, 1.2, 64, seed=18)
    want = np.bincount(to_np(idx), minlength=nbins)
    assert want.max() > 1 << 16
    for merge, hot in ((0, 0), (1, 0), (0, 2)):
        counts = torch.full((nbins,), 5, dtype=dtype, device=dev)
        cfg = B.config(B.BUCKET, block=1024, tile_log2=16, merge=merge, hot=hot)
        assert B.histo(idx, counts, cfg=cfg) == 0
        assert np.array_equal(to_np(counts).astype(np.int64), want + 5), (case, merge, hot)


@pytest.mark.parametrize("nbins,n,tile_log2,dist", [
    ((1 << 21) - 5, 1 << 20, 9, 0),    # 4096 buckets: shallowest pages, mostly short ones
    ((1 << 21) - 5, 1 << 20, 9, 1),
    (1 << 16, 3000, 10, 0),            # one chunk, fewer updates than buckets x depth
    ((1 << 24) + 1, 1 << 22, 16, 1),   # 2^16-bin buckets, packed tile, one short last bucket
    (5_000_011, 1 << 22, 13, 0),       # uniform: the sample switches to the deeper cold pages
])
def test_histo_bucket_pages(dev, orc, nbins, n, tile_log2, dist):
    """The bucketed plan ships LDS buffers as pages of per-chunk pools; short
    pages, the deepest and shallowest page sizes, and the adaptive switch."""
    idx = gen(dev, n, nbins, dist, 1.2, 0, seed=nbins ^ n)
    want = np.zeros(nbins, np.int64)
    orc.histo_agi(to_np(idx), want)
    for hot, merge in ((0, 0), (1, 0), (2, 0), (0, 1)):
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        cfg = B.config(B.BUCKET, block=1024, tile_log2=tile_log2, hot=hot, merge=merge)
        assert B.histo(idx, counts, cfg=cfg) == 0
        assert np.array_equal(to_np(counts), want), (hot, merge)


@pytest.mark.parametrize("nbins,n,dist", [
    (1 << 17, 1 << 21, 0),           # narrowest planned buckets (2^10 bins)
    ((1 << 20) + 3, 1 << 21, 0),     # ragged last bucket, uniform: whole-bucket segments
    ((1 << 20) + 3, 1 << 21, 1),     # skewed: the sketch keeps the shorter segments
    ((1 << 23) + 5, 1 << 22, 0),     # 2^16-bin (packed 16-bit) buckets, fewer updates than bins
    ((1 << 22) + 5, 1 << 22, 1),     # 2^15-bin (32-bit) buckets, skewed
])
def test_histo_segment_choice(dev, orc, nbins, n, dist):
    """The planned bucketed histogram picks its apply segment on the device
    from the sketch (whole buckets without hot bins); a fixed segment length
    (task_log2) and narrow buckets (about 2048 of them) give the same counts."""
    idx = gen(dev, n, nbins, dist, 1.1, 0, seed=nbins + 7 * n)
    want = np.zeros(nbins, np.int64)
    orc.histo_agi(to_np(idx), want)
    plan = B.histo_plan(n, nbins)
    assert plan.strategy == B.BUCKET and plan.hot == 2
    # about 256 buckets, 2^10..2^16 bins each
    assert plan.tile_log2 == max(10, min(16, (nbins - 1).bit_length() - 8))
    cfgs = [plan, B.config(B.BUCKET, block=1024, tile_log2=plan.tile_log2, hot=2, task_log2=14),
            B.config(B.BUCKET, block=1024, tile_log2=max(9, (nbins - 1).bit_length() - 12), hot=2)]
    for cfg in cfgs:
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        assert B.histo(idx, counts, cfg=cfg) == 0
        assert np.array_equal(to_np(counts), want), cfg.as_dict()


@pytest.mark.parametrize("distinct", [300, 1100])
@pytest.mark.parametrize("dtype", [torch.int64, torch.int32])
def test_histo_packed_tile_tally_table(dev, distinct, dtype):
    """What wrapped packed counters are owed is kept in an LDS table of 1024
    bins; one segment of 1100 bins that each wrap (each wrap of a low bin
    also notes its high neighbour) overflows it, and the rest go to device
    atomics (the flush then turns atomic too); 300 such bins fit."""
    nbins, reps = 1 << 16, (1 << 16) + 100
    bins = (torch.arange(distinct, device=dev, dtype=torch.int64) * 59) % nbins
    idx = bins.repeat(reps)
    idx = idx[torch.randperm(idx.numel(), device=dev)]
    want = np.bincount(to_np(idx), minlength=nbins) + 3
    counts = torch.full((nbins,), 3, dtype=dtype, device=dev)
    cfg = B.config(B.BUCKET, block=1024, tile_log2=16, hot=0, task_log2=27)
    assert B.histo(idx, counts, cfg=cfg) == 0
    assert np.array_equal(to_np(counts).astype(np.int64), want)


def test_histo_packed_and_errors(dev, orc):
    nbins, n = 4096, 100_000
    idx = gen(dev, n, nbins, seed=9)
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    want = np.bincount(to_np(idx), minlength=nbins)
    for strat in (B.DIRECT, B.LDS, B.BUCKET):
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        B.histo(pck, counts, packed=True, cfg=B.config(strat, block=1024))
        assert np.array_equal(to_np(counts), want)
    # out-of-range indices are skipped and reported; the rest still count
    bad = idx.clone()
    bad[::1000] = nbins + 5
    bad[1::1000] = -3
    nbad = 2 * len(range(0, n, 1000))
    good = to_np(bad)
    good = good[(good >= 0) & (good < nbins)]
    for strat, hot in ((B.DIRECT, 0), (B.LDS, 0), (B.BUCKET, 0), (B.BUCKET, 2)):
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        with pytest.raises(B.BaleError) as ei:
            B.histo(bad, counts, cfg=B.config(strat, block=1024, hot=hot))
        assert ei.value.code == B.ERR_RANGE
        assert str(nbad) in str(ei.value)
        assert np.array_equal(to_np(counts), np.bincount(good, minlength=nbins))
    # a packed index naming PE 3 cannot run on one GPU
    pck2 = pck.clone()
    pck2[7] |= 3
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    with pytest.raises(B.BaleError) as ei:
        B.histo(pck2, counts, packed=True, cfg=B.config(B.DIRECT))
    assert ei.value.code == B.ERR_PE


def test_histo_empty(dev):
    counts = torch.zeros(10, dtype=torch.int64, device=dev)
    idx = torch.empty(0, dtype=torch.int64, device=dev)
    for strat in (B.DIRECT, B.LDS, B.BUCKET):
        assert B.histo(idx, counts, cfg=B.config(strat)) == 0
    assert int(counts.sum()) == 0


# ---- gather ----------------------------------------------------------------------
@pytest.mark.parametrize("tlen,n,dist,s,run_mean", [
    (1, 100, 0, 1.0, 0),
    (1000, 12345, 0, 1.0, 0),
    (1 << 20, 1 << 20, 0, 1.0, 0),
    (3_000_017, 1 << 20, 1, 1.1, 0),
    (1 << 22, (1 << 20) + 7, 0, 1.0, 64),
])
def test_gather_strategies_match_oracle(dev, orc, tlen, n, dist, s, run_mean):
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    idx = gen(dev, n, tlen, dist, s, run_mean, seed=tlen ^ n)
    want = orc.ig_agi(to_np(idx), orc.fill_splitmix(tlen))
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    for strat, blk, unroll, tile in [(B.DIRECT, 256, 16, 0), (B.DIRECT, 1024, 4, 0), (B.DIRECT, 512, 8, 0),
                                     (B.BUCKET, 1024, 0, 0), (B.BUCKET, 1024, 0, 10), (B.AUTO, 0, 0, 0)]:
        for packed in (False, True):
            tgt = torch.full((n,), 7, dtype=torch.int64, device=dev)
            B.gather(tgt, pck if packed else idx, table, packed=packed,
                     cfg=B.config(strat, block=blk, unroll=unroll, tile_log2=tile))
            assert np.array_equal(to_np(tgt), want), (strat, blk, unroll, tile, packed)


@pytest.mark.parametrize("dist,s,run_mean", [(0, 1.0, 0), (1, 1.1, 0), (1, 1.2, 64), (0, 1.0, 64)],
                         ids=["uniform", "zipf1.1", "zipf_runs", "uniform_runs"])
def test_gather_adaptive_plan(dev, orc, dist, s, run_mean):
    """The planned gather for large shapes is BUCKET with a device-side sample
    that hands inputs repeating entries to DIRECT: either way tgt equals the
    oracle's ig_agi (apps/ig_src/ig_agi.upc:61-64), plain and packed."""
    tlen, n = (1 << 22) + 3, 1 << 24
    plan = B.gather_plan(n, tlen)
    assert plan.strategy == B.BUCKET and plan.hot == 2
    assert B.gather_plan(1 << 23, 1 << 26).strategy == B.DIRECT  # fewer requests than entries / 2
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    idx = gen(dev, n, tlen, dist, s, run_mean, seed=31 + dist + run_mean)
    want = orc.ig_agi(to_np(idx), orc.fill_splitmix(tlen))
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    for packed, src in ((False, idx), (True, pck)):
        tgt = torch.full((n,), 5, dtype=torch.int64, device=dev)
        B.gather(tgt, src, table, packed=packed)
        assert B.lib().bale_hip_last_strategy() == B.BUCKET
        assert np.array_equal(to_np(tgt), want), packed


@pytest.mark.parametrize("tlen,want_w", [((1 << 22) + 3, 13), ((1 << 24) - 1, 13), ((1 << 25) + 7, 14),
                                         ((1 << 26) + 1, 18)])
def test_gather_plan_regions(dev, orc, tlen, want_w):
    """The planned BUCKET gather serves tables of up to 2^25 entries from LDS
    regions of 2^13-2^14 entries (one workgroup per region, its slice of the
    table copied in), larger ones from 2^18-entry L2 regions; ragged tables end
    in a short region.  Plain and packed requests equal the oracle's ig_agi."""
    n = max(1 << 24, (tlen + 1) // 2)  # the plan's BUCKET needs a request per two entries
    plan = B.gather_plan(n, tlen)
    assert plan.strategy == B.BUCKET and plan.tile_log2 == want_w
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    idx = gen(dev, n, tlen, seed=tlen & 0xffff)
    idx[-1] = tlen - 1  # the last entry of the short last region
    want = orc.ig_agi(to_np(idx), orc.fill_splitmix(tlen))
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    for cfg in (plan, B.config(B.BUCKET, tile_log2=want_w, hot=0)):
        for packed, src in ((False, idx), (True, pck)):
            tgt = torch.full((n,), 5, dtype=torch.int64, device=dev)
            B.gather(tgt, src, table, packed=packed, cfg=cfg)
            assert np.array_equal(to_np(tgt), want), (cfg, packed)


@pytest.mark.parametrize("tlen,w", [((1 << 25) - 5, 13), ((1 << 26) - 9, 14), (4095 * (1 << 11) + 1, 11)])
def test_gather_most_regions(dev, orc, tlen, w):
    """At the most regions a chunk's sort counts (4096, the last one short)
    every chunk's run cursors fill the sort's LDS counter array: plain and
    packed requests, with a ragged last chunk, equal the oracle's ig_agi."""
    assert (tlen + (1 << w) - 1) >> w == 4096
    n = (1 << 22) + 333
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    idx = gen(dev, n, tlen, seed=w)
    idx[0], idx[-1] = 0, tlen - 1
    want = orc.ig_agi(to_np(idx), orc.fill_splitmix(tlen))
    pck = torch.empty_like(idx)
    B.pack(pck, idx)
    for packed, src in ((False, idx), (True, pck)):
        tgt = torch.full((n,), 5, dtype=torch.int64, device=dev)
        B.gather(tgt, src, table, packed=packed, cfg=B.config(B.BUCKET, tile_log2=w, hot=0))
        assert B.lib().bale_hip_last_strategy() == B.BUCKET
        assert np.array_equal(to_np(tgt), want), packed


def test_gather_bucket_errors_and_ragged(dev, orc):
    """BUCKET gather: out-of-range and foreign-PE requests are counted and
    their slots read 0; a ragged last chunk; regions narrower than default."""
    tlen, n = 100_003, 3 * (1 << 14) + 77
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_splitmix(table)
    idx = gen(dev, n, tlen, seed=12)
    bad = torch.arange(0, n, 97, device=dev)
    idx[bad] = tlen + 11
    want = orc.ig_agi(np.minimum(to_np(idx), tlen - 1), orc.fill_splitmix(tlen))
    want[to_np(bad)] = 0
    for tile in (0, 10, 12):
        tgt = torch.full((n,), 5, dtype=torch.int64, device=dev)
        with pytest.raises(B.BaleError) as e:
            B.gather(tgt, idx, table, cfg=B.config(B.BUCKET, tile_log2=tile))
        assert e.value.code == B.ERR_RANGE
        assert np.array_equal(to_np(tgt), want), tile
    pck = torch.empty_like(idx)
    B.pack(pck, idx.clamp(max=tlen - 1))
    pck[7] += 3  # PE 3
    tgt = torch.full((n,), 5, dtype=torch.int64, device=dev)
    with pytest.raises(B.BaleError) as e:
        B.gather(tgt, pck, table, packed=True, cfg=B.config(B.BUCKET))
    assert e.value.code == B.ERR_PE
    w2 = orc.ig_agi(np.minimum(to_np(idx), tlen - 1), orc.fill_splitmix(tlen))
    w2[7] = 0
    assert np.array_equal(to_np(tgt), w2)


def test_gather_closed_form_reference_table(dev, orc):
    tlen, n = 100_000, 1_000_000
    table = torch.empty(tlen, dtype=torch.int64, device=dev)
    B.fill_neg(table)  # ig.upc:159-161
    idx = torch.from_numpy(orc.reference_rand(n, tlen, 5)).to(dev)  # ig.upc:168-170
    tgt = torch.zeros(n, dtype=torch.int64, device=dev)
    B.gather(tgt, idx, table)
    assert torch.equal(tgt, -(idx + 1))  # ig_check_and_zero, ig.upc:96-114


def test_gather_range_error(dev):
    table = torch.arange(100, dtype=torch.int64, device=dev)
    idx = torch.tensor([0, 5, 100, 99, -1], dtype=torch.int64, device=dev)
    tgt = torch.zeros(5, dtype=torch.int64, device=dev)
    with pytest.raises(B.BaleError) as ei:
        B.gather(tgt, idx, table)
    assert ei.value.code == B.ERR_RANGE
    t = to_np(tgt)
    assert t[0] == 0 and t[1] == 5 and t[3] == 99


# ---- reference drop-ins ------------------------------------------------------------
def ptr(a):
    return a.ctypes.data if isinstance(a, np.ndarray) else a.data_ptr()


@pytest.mark.parametrize("where", ["host", "device"])
def test_histo_dropins_reference_check(dev, orc, where):
    """The reference driver's protocol (histo.upc:161-227): every model adds into
    the same counts; subtracting num_models per index leaves zeros."""
    L = B.lib()
    nbins, n = 1000, 1_000_000  # the reference defaults -T 1000 -n 1000000 at one PE
    idx_h = orc.reference_rand(n, nbins, 120348)
    pck_h = orc.pack(idx_h, 1)
    if where == "host":
        idx, pck, counts = idx_h, pck_h, np.zeros(nbins, np.int64)
    else:
        idx, pck = torch.from_numpy(idx_h).to(dev), torch.from_numpy(pck_h).to(dev)
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    models = 0
    assert L.histo_agi(ptr(idx), n, ptr(counts)) >= 0, B.last_error()
    models += 1
    for name in B.DROPIN_HISTO_BUF:
        t = getattr(L, name)(ptr(pck), n, ptr(counts), 1024)
        assert t >= 0, (name, B.last_error())
        models += 1
    assert L.histo_conveyor(ptr(pck), n, ptr(counts)) >= 0, B.last_error()
    models += 1
    c = counts if where == "host" else to_np(counts)
    want = np.bincount(idx_h, minlength=nbins) * models
    assert np.array_equal(c, want)
    np.subtract.at(c, idx_h, models)
    assert not c.any()


@pytest.mark.parametrize("where", ["host", "device"])
def test_ig_dropins_closed_form(dev, orc, where):
    L = B.lib()
    tlen, n = 100_000, 1_000_000  # the reference defaults -T 100000 -n 1000000
    table_h = orc.fill_neg(tlen)
    idx_h = orc.reference_rand(n, tlen, 5)
    pck_h = orc.pack(idx_h, 1)
    if where == "host":
        table, idx, pck = table_h, idx_h, pck_h
        mk = lambda: np.zeros(n, np.int64)  # noqa: E731
        get = lambda t: t  # noqa: E731
    else:
        table, idx, pck = (torch.from_numpy(a).to(dev) for a in (table_h, idx_h, pck_h))
        mk = lambda: torch.zeros(n, dtype=torch.int64, device=dev)  # noqa: E731
        get = to_np
    want = -(idx_h + 1)
    tgt = mk()
    assert L.ig_agi(ptr(tgt), ptr(idx), n, ptr(table)) >= 0, B.last_error()
    assert np.array_equal(get(tgt), want)
    for name in B.DROPIN_IG_BUF + ["ig_conveyor"]:
        tgt = mk()
        fn = getattr(L, name)
        t = fn(ptr(tgt), ptr(pck), n, ptr(table)) if name == "ig_conveyor" else \
            fn(ptr(tgt), ptr(pck), n, ptr(table), 1024)
        assert t >= 0, (name, B.last_error())
        assert np.array_equal(get(tgt), want), name


def test_dropin_timing_and_errors(dev, orc):
    L = B.lib()
    idx = np.array([0, 1, 2, -4], np.int64)
    counts = np.zeros(8, np.int64)
    assert L.histo_agi(ptr(idx), 4, ptr(counts)) == -1.0
    assert "negative" in B.last_error()
    pck = np.array([(1 << 16) | 1], np.int64)
    assert L.histo_exstack(ptr(pck), 1, ptr(counts), 1024) == -1.0
    assert "PE" in B.last_error()
    idx = orc.gen_uniform(1 << 20, 1 << 16, 1)
    counts = np.zeros(1 << 16, np.int64)
    t = L.histo_agi(ptr(idx), idx.size, ptr(counts))
    g, h2d, d2h = B.last_timing()
    assert t == g and t > 0 and h2d > 0 and d2h > 0
