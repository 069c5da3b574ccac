"""The library boundary on the GPU: configurations a shape cannot take, one
error record and workspace per stream, the drop-ins' strategy choice for large
tables, and batches whose histogram instances share one counts table.

Results are checked against the oracle's histo_agi / ig_agi
(apps/histo_src/histo_agi.upc:61-66, apps/ig_src/ig_agi.upc:61-64)."""
import ctypes as C

import numpy as np
import pytest
import torch

import bale_old_amd as B
from bale_old_amd import mix as M

pytestmark = pytest.mark.gpu


def to_np(t):
    return t.cpu().numpy()


def gen(dev, n, nbins, dist=0, s=1.0, run_mean=0, seed=1):
    out = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(out, nbins, dist=dist, s=s, run_mean=run_mean, seed=seed)
    return out


@pytest.mark.parametrize("strat,tile,nbins,words", [
    (B.BUCKET, 9, 1 << 24, "tile_log2 >= 12"),      # 32768 buckets
    (B.BUCKET, 16, (1 << 28) + 1, "use DIRECT"),     # more than 4096 buckets of 2^16 bins
    (B.LDS, 8, 1 << 26, "65535"),                    # 262144 bin tiles
])
def test_shapes_a_config_cannot_take_are_argument_errors(dev, strat, tile, nbins, words):
    idx = torch.zeros(4096, dtype=torch.int64, device=dev)
    counts = torch.zeros(1, dtype=torch.int64, device=dev)  # never touched: refused first
    cfg = B.config(strat, block=1024, tile_log2=tile)
    rc = B.lib().bale_histo(B._ptr(idx), idx.numel(), B._ptr(counts), nbins, B.I64, B.PLAIN,
                            C.byref(cfg), None)
    assert rc == B.ERR_ARG, B.last_error()
    assert words in B.last_error()


def test_bucket_groups_per_cu_that_leave_no_lds(dev):
    nbins = 1 << 22
    idx = gen(dev, 1 << 20, nbins, seed=3)
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    cfg = B.config(B.BUCKET, block=1024, tile_log2=10, groups_per_cu=16)
    rc = B.lib().bale_histo(B._ptr(idx), idx.numel(), B._ptr(counts), nbins, B.I64, B.PLAIN,
                            C.byref(cfg), None)
    assert rc in (B.OK, B.ERR_ARG)
    if rc == B.ERR_ARG:
        assert "LDS" in B.last_error() or "16-byte" in B.last_error()
    else:
        B.sync_errors()
        assert np.array_equal(to_np(counts), np.bincount(to_np(idx), minlength=nbins))


def test_streams_have_their_own_workspace_and_errors(dev, orc):
    """Bucketed histograms of different sizes in flight on two streams at once
    (each grows its own workspace), and an out-of-range index on one stream
    is reported there and not on the other."""
    s1, s2 = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)
    cases = []
    for k, (nbins, n) in enumerate([(1 << 20, 1 << 22), (1 << 22, 1 << 23), (3 << 18, 1 << 21)]):
        idx = gen(dev, n, nbins, dist=k % 2, s=1.1, seed=40 + k)
        want = np.zeros(nbins, np.int64)
        orc.histo_agi(to_np(idx), want)
        cases.append((idx, nbins, want))
    torch.cuda.synchronize()
    outs = []
    for rep in range(3):
        for k, (idx, nbins, want) in enumerate(cases):
            st = s1 if (k + rep) % 2 == 0 else s2
            counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
            st.wait_stream(torch.cuda.current_stream())  # the zero fill first
            plan = B.histo_plan(idx.numel(), nbins)
            B.histo(idx, counts, cfg=plan, stream=st, sync=False)
            outs.append((counts, want))
    bad_idx = torch.tensor([0, 5, 1 << 40, 7], dtype=torch.int64, device=dev)
    c_bad = torch.zeros(16, dtype=torch.int64, device=dev)
    s1.wait_stream(torch.cuda.current_stream())
    B.histo(bad_idx, c_bad, cfg=B.config(B.DIRECT), stream=s1, sync=False)
    assert B.sync_errors(s2) == 0
    with pytest.raises(B.BaleError) as e:
        B.sync_errors(s1)
    assert e.value.code == B.ERR_RANGE
    assert B.sync_errors(s1) == 0  # cleared
    for counts, want in outs:
        assert np.array_equal(to_np(counts), want)
    assert to_np(c_bad).tolist() == [1] + [0] * 4 + [1] + [0] + [1] + [0] * 8
    for st in (s1, s2):
        assert B.lib().bale_hip_release_stream(C.c_void_p(st.cuda_stream)) == B.OK
    # a released stream starts again with a fresh record and workspace
    counts = torch.zeros(cases[0][1], dtype=torch.int64, device=dev)
    B.histo(cases[0][0], counts, stream=s1)
    assert np.array_equal(to_np(counts), cases[0][2])


def test_dropin_lds_models_fall_back_to_the_plan_for_large_tables(dev):
    """histo_exstack2 and histo_exstack2_function take the LDS strategy, which
    reads the index once per bin tile: up to one tile per CU they run it, for
    larger tables the library's plan."""
    L = B.lib()
    cus, _ = B.device_info()
    for nbins, expect_lds in ((1 << 22, True), (1 << 24, False)):
        n = 1 << 22
        idx = gen(dev, n, nbins, seed=nbins)
        pck = torch.empty_like(idx)
        B.pack(pck, idx)
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        for name in ("histo_exstack2", "histo_exstack2_function", "histo_exstack2_cyclic"):
            t = getattr(L, name)(B._ptr(pck), n, B._ptr(counts), 1024)
            assert t >= 0, B.last_error()
            took = L.bale_hip_last_strategy()
            tiles = -(-nbins // (1 << 15))
            assert (took == B.LDS) == (expect_lds and tiles <= cus), (name, nbins, took)
            if not expect_lds:
                assert took == B.BUCKET
        want = 3 * np.bincount(to_np(idx), minlength=nbins)
        assert np.array_equal(to_np(counts), want)


def test_batch_histograms_sharing_one_table(dev, orc):
    """Several index streams into one counts table in one batch: bucketed
    instances (whose single-segment buckets would flush with plain adds),
    a sub-range instance, LDS and direct ones, int64 and int32."""
    for dtype, bdt in ((torch.int64, B.I64), (torch.int32, B.I32)):
        nbins = 1 << 20
        counts = torch.full((nbins,), 11, dtype=dtype, device=dev)
        # (updates, first bin, bins, distribution, configuration)
        specs = [(1 << 21, 0, nbins, 0, B.config(B.BUCKET, hot=2)),
                 (1 << 21, 0, nbins, 1, B.config(B.BUCKET, hot=0)),
                 (1 << 20, 0, nbins, 0, B.config(B.BUCKET, hot=1, merge=1)),
                 (1 << 19, 1 << 18, 1 << 16, 0, B.config(B.BUCKET, tile_log2=10)),
                 (1 << 18, 0, nbins, 0, B.config(B.DIRECT)),
                 (1 << 18, (1 << 18) + 4096, 4096, 1, B.config(B.LDS))]
        insts, idxs = [], []
        want = np.full(nbins, 11, np.int64)
        for k, (n, off, tl, dist, cfg) in enumerate(specs):
            idx = gen(dev, n, tl, dist=dist, s=1.1, seed=70 + k)
            idxs.append(idx)
            w = np.zeros(tl, np.int64)
            orc.histo_agi(to_np(idx), w)
            want[off:off + tl] += w
            inst = B.Instance()
            inst.kind, inst.dtype, inst.n, inst.table_len = 0, bdt, n, tl
            inst.index = B._ptr(idx)
            inst.table = B._ptr(counts) + off * counts.element_size()
            inst.tgt = None
            inst.cfg = cfg
            insts.append(inst)
        arr = (B.Instance * len(insts))(*insts)
        torch.cuda.synchronize()
        b = M.Batch(arr, len(insts))
        for _ in range(2):
            b.launch()
        B.sync_errors()
        want2 = want + (want - 11)  # two launches add twice
        assert np.array_equal(to_np(counts).astype(np.int64), want2), dtype
        b.close()
