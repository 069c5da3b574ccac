"""exstack and exstack2 on the GPU (bale_exstack*, bale_exstack2*) against the
oracle.

Per-destination FIFO order is compared with oracle.convey_deliver (pinned in
tests/test_oracle.py by a stable sort and per-destination queues); the
(destination, source) grouping with numpy's stable sort; the request/reply
retrace of apps/ig_src/ig_exstack.upc:69-94 run through the API with
oracle.ig_exstack; the histogram loops of apps/histo_src/histo_exstack.upc:64-76
and histo_exstack2.upc:63-74 with oracle.histo_agi."""
import numpy as np
import pytest
import torch

import bale_old_amd as B
from bale_old_amd.exstack import Exstack

pytestmark = pytest.mark.gpu


def rand_items(n, ib, P, seed):
    g = torch.Generator().manual_seed(seed)
    items = torch.randint(0, 256, (n, ib), dtype=torch.uint8, generator=g)
    to = torch.randint(0, P, (n,), dtype=torch.int64, generator=g)
    frm = torch.randint(0, P, (n,), dtype=torch.int64, generator=g)
    return items, to, frm


@pytest.mark.parametrize("P,ib,n,buf", [(1, 8, 1000, 64), (7, 8, 100_003, 1024), (255, 4, 300_000, 512),
                                        (64, 16, 65_536, 100), (13, 12, 4097, 7), (3, 32, 20_000, 10_000)])
def test_exstack_fifo_per_destination(dev, orc, P, ib, n, buf):
    """Pushing from PE 0 round by round: every destination receives its items
    in push order, over all rounds the oracle's FIFO delivery."""
    items, to, _ = rand_items(n, ib, P, seed=P * 7 + n)
    it, tp = items.cuda(), to.cuda()
    ex = Exstack(P, buf, ib)
    got = [[] for _ in range(P)]
    i, rounds = 0, 0
    while ex.proceed(done=(i == n)):
        k = ex.push(it[i:], tp[i:])
        assert k == min(n - i, P * buf)
        i += k
        ex.exchange()
        rounds += 1
        for p in range(P):
            a, frm = ex.pop(p)
            assert int(frm.abs().sum()) == 0  # one source: PE 0
            got[p].append(a.cpu().numpy())
            assert ex.pop(p)[0].shape[0] == 0  # popped
    assert rounds == -(-n // (P * buf))
    want, starts, dropped = orc.convey_deliver(items.numpy(), to.numpy(), P)
    assert dropped == 0
    for p in range(P):
        assert np.array_equal(np.concatenate(got[p]), want[starts[p]:starts[p + 1]]), p
    ex.clear()


@pytest.mark.parametrize("P,n", [(5, 50_000), (64, 200_000), (255, 100_000)])
def test_exstack_groups_by_source(dev, P, n):
    """With sources named, a destination's round comes grouped by source in
    increasing order, each source's items in push order; pop_thread hands out
    one source's group."""
    items, to, frm = rand_items(n, 8, P, seed=n + P)
    ex = Exstack(P, n, 8)
    assert ex.push(items.cuda(), to.cuda(), frm.cuda()) == n
    ex.exchange()
    it, t, f = items.numpy(), to.numpy(), frm.numpy()
    order = np.lexsort((np.arange(n), f, t))  # by destination, then source, then push order
    for p in range(0, P - 1, max(1, P // 9)):  # the last PE is popped source by source below
        a, s = ex.pop(p)
        sel = order[t[order] == p]
        assert np.array_equal(a.cpu().numpy(), it[sel]), p
        assert np.array_equal(s.cpu().numpy(), f[sel]), p
    q = P - 1
    for src in range(P):
        a = ex.pop_thread(q, src)
        sel = np.nonzero((t == q) & (f == src))[0]
        assert np.array_equal(a.cpu().numpy(), it[sel]), src
    with pytest.raises(B.BaleError):
        ex.pop(q)  # popped one source at a time this round
    ex.clear()


@pytest.mark.parametrize("P", [1, 7, 16])
def test_ig_exstack_retrace_through_the_api(dev, orc, P):
    """apps/ig_src/ig_exstack.upc:69-94 on the API: requests (local indices)
    go to their owners, each owner answers its requests in place and pushes
    the replies back to the source, and the source retraces its round's
    requests in order, popping each owner's replies (pop_thread).  tgt equals
    the oracle's ig_exstack on the same packed indices."""
    tab_len, n, buf = 9973, 60_000, 1000
    table = orc.fill_splitmix(tab_len)
    idx = orc.gen_index(n, tab_len, 1, 1.2, 8, 6)
    lnum = (tab_len + P - 1) // P
    ltables = np.zeros(lnum * P, np.int64)
    for pe in range(P):
        js = np.arange(lnum) * P + pe
        ok = js < tab_len
        ltables[pe * lnum:(pe + 1) * lnum][ok] = table[js[ok]]
    pck = orc.pack(idx, P)
    want = orc.ig_exstack(pck, ltables, P)
    pck_d = torch.from_numpy(pck).to(dev)
    lidx_d, pe_d = pck_d >> 16, pck_d & 0xFFFF
    lt_d = torch.from_numpy(ltables).to(dev).view(P, lnum)
    zero = torch.zeros(1, dtype=torch.int64, device=dev)
    ex = Exstack(P, buf, 8)
    tgt = torch.empty(n, dtype=torch.int64, device=dev)
    i = 0
    while ex.proceed(done=(i == n)):
        i0 = i
        i += ex.push(lidx_d[i:].contiguous(), pe_d[i:].contiguous())  # from PE 0 (the requester)
        ex.exchange()
        # owners: pop their requests, answer in place, push the replies back to the source
        for p in range(P):
            req, frm = ex.pop(p)
            if req.shape[0]:
                rep = lt_d[p][req.view(torch.int64).reshape(-1)]
                back = torch.full_like(frm, p)
                assert ex.push(rep.contiguous(), frm.contiguous(), back) == rep.numel()
        ex.exchange()
        # retrace: request j's reply is the next one from its owner
        pes = pe_d[i0:i]
        rank = torch.zeros_like(pes)
        for p in range(P):
            m = pes == p
            rank[m] = torch.arange(int(m.sum()), device=dev)
        for p in range(P):
            reps = ex.pop_thread(0, p).view(torch.int64).reshape(-1)
            m = pes == p
            tgt[i0:i][m] = reps[rank[m]]
    assert np.array_equal(tgt.cpu().numpy(), want)
    ex.clear()


@pytest.mark.parametrize("kind", [1, 2])
def test_histogram_loops(dev, orc, kind):
    """histo_exstack (exchange per round) and histo_exstack2 (proceed sends)
    on packed indices: every PE adds what it received to its counts."""
    P, n, nbins, buf = 8, 200_000, 50_000, 2048
    idx = orc.gen_index(n, nbins, 1, 1.1, 0, 3)
    pck = torch.from_numpy(orc.pack(idx, P)).to(dev)
    col, pe = (pck >> 16).contiguous(), (pck & 0xFFFF).contiguous()
    lnum = (nbins + P - 1) // P
    lcounts = torch.zeros((P, lnum), dtype=torch.int64, device=dev)
    ex = Exstack(P, buf, 8, kind=kind)
    i = 0
    if kind == 1:
        while ex.proceed(done=(i == n)):
            if i < n:
                i += ex.push(col[i:], pe[i:])
            ex.exchange()
            for p in range(P):
                got, _ = ex.pop(p)
                lcounts[p].index_add_(0, got.view(torch.int64).reshape(-1),
                                      torch.ones(got.shape[0], dtype=torch.int64, device=dev))
    else:
        while ex.proceed(done=(i == n)):
            if i < n:
                i += ex.push(col[i:], pe[i:])
            for p in range(P):
                got, _ = ex.pop(p)
                lcounts[p].index_add_(0, got.view(torch.int64).reshape(-1),
                                      torch.ones(got.shape[0], dtype=torch.int64, device=dev))
    want = np.zeros(nbins, np.int64)
    orc.histo_agi(idx, want)
    glob = lcounts.t().reshape(-1)[:nbins].cpu().numpy()  # global j = l * P + pe
    assert np.array_equal(glob, want)
    ex.clear()


def test_bad_pes_are_dropped_and_reported(dev):
    P, n = 6, 10_000
    items, to, frm = rand_items(n, 4, P, seed=5)
    to[::97] = P + 2
    frm[5::101] = -1
    ex = Exstack(P, n, 4)
    assert ex.push(items.cuda(), to.cuda(), frm.cuda()) == n
    with pytest.raises(B.BaleError) as e:
        ex.exchange()
    assert e.value.code == B.ERR_RANGE
    ok = (to < P) & (frm >= 0)
    total = 0
    for p in range(P):
        a, s = ex.pop(p)
        sel = np.nonzero((ok & (to == p)).numpy())[0]
        order = sel[np.argsort(frm.numpy()[sel], kind="stable")]
        assert np.array_equal(a.cpu().numpy(), items.numpy()[order])
        total += a.shape[0]
    assert total == int(ok.sum())
    assert not ex.proceed(done=True)
    ex.clear()


def test_argument_errors(dev):
    with pytest.raises(B.BaleError):
        Exstack(256, 10, 8)
    with pytest.raises(B.BaleError):
        Exstack(4, 10, 10)
    ex = Exstack(4, 10, 8, kind=2)
    ex.proceed(done=True)
    with pytest.raises(B.BaleError):
        ex.push(torch.zeros(1, dtype=torch.int64, device=dev), torch.zeros(1, dtype=torch.int64, device=dev))
    ex.clear()
