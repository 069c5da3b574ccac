"""bale_histo_noalloc (the entry point that allocates nothing) against
bale_histo through the C-ABI: the same counts, on two calls in a row (the
second smaller, so the shared scratch holds the first call's leftovers), for
the bucketed plan (packed and 32-bit tiles, with and without the hot-bin
sample), DIRECT with the sample, and LDS tiles; sizes beyond the scratch are
refused; index errors go to the caller's counters.  Both sides are checked
against the oracle's histo_agi (apps/histo_src/histo_agi.upc:61-66)."""
import ctypes as C

import numpy as np
import pytest
import torch

import bale_old_amd as B

pytestmark = pytest.mark.gpu


def gen(dev, n, nbins, dist=0, s=1.0, seed=1):
    out = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(out, nbins, dist=dist, s=s, seed=seed)
    return out


def noalloc(idx, counts, cfg, errors=None, packed=False):
    dtype = B.I64 if counts.dtype == torch.int64 else B.I32
    return B.lib().bale_histo_noalloc(
        B._ptr(idx), idx.numel(), B._ptr(counts), counts.numel(), dtype,
        B.PACKED if packed else B.PLAIN, C.byref(cfg) if cfg is not None else None,
        B._ptr(errors), None)


CASES = [
    # (bins, first call's updates, second call's updates, distribution, config)
    (1 << 20, 1 << 24, 1 << 21, 0, None),                                   # the plan: BUCKET, hot = 2
    (1 << 20, 1 << 23, 3 << 19, 1, None),                                   # same, skewed input
    ((1 << 22) + 5, 1 << 23, 1 << 20, 1, B.config(B.BUCKET, block=1024, tile_log2=16, hot=1)),
    (100_003, 1 << 22, 12345, 0, B.config(B.BUCKET, block=1024, tile_log2=12, hot=0, merge=1)),
    (1 << 18, 1 << 22, 1 << 19, 1, B.config(B.DIRECT, hot=1)),
    (1 << 15, 1 << 22, 77777, 0, B.config(B.LDS, block=1024)),
]


@pytest.mark.parametrize("nbins,n1,n2,dist,cfg", CASES)
@pytest.mark.parametrize("dtype", [torch.int64, torch.int32])
def test_noalloc_matches_bale_histo(dev, nbins, n1, n2, dist, cfg, dtype):
    for k, n in enumerate((n1, n2)):
        idx = gen(dev, n, nbins, dist=dist, s=1.1, seed=nbins + n + k)
        base = torch.randint(0, 100, (nbins,), dtype=dtype, device=dev)
        got, ref = base.clone(), base.clone()
        torch.cuda.synchronize()
        assert noalloc(idx, got, cfg) == B.OK, B.last_error()
        torch.cuda.synchronize()
        assert B.lib().bale_histo(B._ptr(idx), n, B._ptr(ref), nbins,
                                  B.I64 if dtype == torch.int64 else B.I32, B.PLAIN,
                                  C.byref(cfg) if cfg is not None else None, None) == B.OK
        B.sync_errors()
        g = got.cpu().numpy().astype(np.int64)
        assert np.array_equal(g, ref.cpu().numpy().astype(np.int64)), (nbins, n, k)
        want = base.cpu().numpy().astype(np.int64) + np.bincount(idx.cpu().numpy(), minlength=nbins)
        assert np.array_equal(g, want), (nbins, n, k)


def test_noalloc_refuses_more_than_its_scratch(dev):
    nbins, n = 1 << 24, 1 << 26
    idx = gen(dev, n, nbins, seed=5)
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    rc = noalloc(idx, counts, B.config(B.BUCKET, block=1024, hot=2, tile_log2=16))
    assert rc == B.ERR_ARG
    assert "bytes of scratch" in B.last_error()
    torch.cuda.synchronize()
    assert int(counts.count_nonzero()) == 0
    # the same shape through DIRECT needs no scratch
    assert noalloc(idx, counts, B.config(B.DIRECT)) == B.OK
    torch.cuda.synchronize()
    assert int(counts.sum()) == n


def test_noalloc_reports_errors_to_the_callers_counters(dev):
    nbins = 1000
    idx = gen(dev, 1 << 16, nbins, seed=6)
    idx[::1000] = nbins + 3
    bad = int((idx >= nbins).sum())
    errors = torch.zeros(2, dtype=torch.int64, device=dev)
    for cfg in (B.config(B.DIRECT), B.config(B.LDS, block=1024), B.config(B.BUCKET, block=1024)):
        counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
        errors.zero_()
        torch.cuda.synchronize()
        assert noalloc(idx, counts, cfg, errors) == B.OK
        torch.cuda.synchronize()
        assert errors.cpu().tolist() == [bad, 0], cfg.as_dict()
        ok = idx[idx < nbins].cpu().numpy()
        assert np.array_equal(counts.cpu().numpy(), np.bincount(ok, minlength=nbins))
    # packed indices naming PE 1
    pck = torch.empty_like(idx)
    B.pack(pck, idx.clamp(max=nbins - 1))
    pck[5] += 1
    errors.zero_()
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    assert noalloc(pck, counts, None, errors, packed=True) == B.OK
    torch.cuda.synchronize()
    assert errors.cpu().tolist() == [0, 1]
