"""Device conveyor (bale_convey_*) against the oracle's FIFO delivery.

The reference pins conveyors with its own send/receive checksums
(convey/alltoallv.c); here every PE's delivered span is compared item for
item with oracle.convey_deliver, which tests/test_oracle.py pins against a
stable sort and per-destination FIFO queues."""
import numpy as np
import pytest
import torch

import bale_old_amd as B
from bale_old_amd.convey import DONE, OK, Conveyor

pytestmark = pytest.mark.gpu


def rand_items(n, ib, P, seed, lo=0, hi=None):
    g = torch.Generator().manual_seed(seed)
    items = torch.randint(0, 256, (n, ib), dtype=torch.uint8, generator=g)
    pes = torch.randint(lo, P if hi is None else hi, (n,), dtype=torch.int64, generator=g)
    return items, pes


@pytest.mark.parametrize("P,ib,n", [(1, 8, 1000), (7, 8, 100_003), (256, 4, 300_000),
                                    (64, 16, 65_536), (13, 12, 4097), (200, 24, 50_000),
                                    (3, 32, 20_000), (2, 8, 1)])
def test_delivery_matches_oracle(dev, orc, P, ib, n):
    items, pes = rand_items(n, ib, P, seed=P * 1000 + ib + n)
    c = Conveyor(P, n, ib)
    c.begin()
    assert c.push(items.cuda(), pes.cuda())
    assert c.advance(True) == OK
    want, starts, dropped = orc.convey_deliver(items.numpy(), pes.numpy(), P)
    assert dropped == 0
    for p in range(P):
        got = c.pull(p).cpu().numpy()
        assert np.array_equal(got, want[starts[p]:starts[p + 1]]), p
    assert c.advance(True) == DONE
    c.free()


def test_several_pushes_one_delivery_and_capacity(dev, orc):
    P, ib = 5, 8
    items, pes = rand_items(30_000, ib, P, seed=7)
    c = Conveyor(P, 20_000, ib)
    c.begin()
    assert c.push(items[:12_000].cuda(), pes[:12_000].cuda())
    assert c.push(items[12_000:20_000].cuda(), pes[12_000:20_000].cuda())
    assert not c.push(items[20_000:].cuda(), pes[20_000:].cuda())  # full: nothing copied
    assert c.advance(False) == OK
    want, starts, _ = orc.convey_deliver(items[:20_000].numpy(), pes[:20_000].numpy(), P)
    # a delivery waits until the previous one is pulled
    assert c.push(items[20_000:].cuda(), pes[20_000:].cuda())
    assert c.advance(True) == OK
    for p in range(P):
        assert np.array_equal(c.pull(p).cpu().numpy(), want[starts[p]:starts[p + 1]])
    assert c.advance(True) == OK  # delivers the rest
    want2, starts2, _ = orc.convey_deliver(items[20_000:].numpy(), pes[20_000:].numpy(), P)
    for p in range(P):
        assert np.array_equal(c.pull(p).cpu().numpy(), want2[starts2[p]:starts2[p + 1]])
    assert c.advance(True) == DONE
    c.free()


def test_histogram_through_the_conveyor(dev, orc):
    """histo_conveyor's pattern (apps/histo_src/histo_conveyor.upc:72-82) over
    8 virtual PEs owning bins cyclically: push the local index to the owner,
    pull and count; the loop runs several deliveries (capacity < updates)."""
    P, nbins, n = 8, 10_007, 200_000
    idx = torch.from_numpy(orc.gen_index(n, nbins, 1, 1.2, 0, 5))
    local = (idx // P).cuda()
    owner = (idx % P).cuda()
    lcounts = [torch.zeros((nbins + P - 1) // P, dtype=torch.int64, device="cuda") for _ in range(P)]
    c = Conveyor(P, 70_000, 8)
    c.begin()
    i = 0
    rounds = 0
    while c.advance(i == n) != DONE:
        rounds += 1
        while i < n:
            j = min(n, i + 25_000)
            if not c.push(local[i:j], owner[i:j]):
                break
            i = j
        for p in range(P):
            got = c.pull(p).view(torch.int64).reshape(-1)
            lcounts[p] += torch.bincount(got, minlength=lcounts[p].numel())
    assert rounds > 2
    counts = np.zeros(nbins, dtype=np.int64)
    for p in range(P):
        counts[p::P] = lcounts[p].cpu().numpy()[: len(range(p, nbins, P))]
    assert np.array_equal(counts, np.bincount(idx.numpy(), minlength=nbins))
    c.free()


def test_pe_out_of_range_is_reported_and_dropped(dev, orc):
    P, ib, n = 6, 8, 10_000
    items, pes = rand_items(n, ib, P, seed=11, lo=-1, hi=P + 2)
    c = Conveyor(P, n, ib)
    c.begin()
    assert c.push(items.cuda(), pes.cuda())
    with pytest.raises(B.BaleError) as ei:
        c.advance(True)
    assert ei.value.code == B.ERR_RANGE
    want, starts, dropped = orc.convey_deliver(items.numpy(), pes.numpy(), P)
    assert dropped > 0 and str(dropped) in str(ei.value)
    for p in range(P):
        assert np.array_equal(c.pull(p).cpu().numpy(), want[starts[p]:starts[p + 1]])
    assert c.advance(True) == DONE
    c.free()


def test_unpull_and_reset(dev, orc):
    P, ib, n = 3, 4, 999
    items, pes = rand_items(n, ib, P, seed=5)
    c = Conveyor(P, n, 8)  # room for 8-byte items; this round uses 4
    c.begin(ib)
    c.push(items.cuda(), pes.cuda())
    c.advance(True)
    want, starts, _ = orc.convey_deliver(items.numpy(), pes.numpy(), P)
    first = c.pull(1).cpu().numpy()
    assert np.array_equal(first, want[starts[1]:starts[2]])
    assert c.unpull(1, 5)
    assert not c.unpull(1, 1000)
    assert np.array_equal(c.pull(1).cpu().numpy(), first[-5:])
    assert c.pull(1).shape[0] == 0
    c.reset()
    c.begin(8)
    items8, pes8 = rand_items(100, 8, P, seed=6)
    c.push(items8.cuda(), pes8.cuda())
    assert c.advance(True) == OK
    want8, st8, _ = orc.convey_deliver(items8.numpy(), pes8.numpy(), P)
    for p in range(P):
        assert np.array_equal(c.pull(p).cpu().numpy(), want8[st8[p]:st8[p + 1]])
    assert c.advance(True) == DONE
    c.free()


def test_large_delivery_keeps_push_order(dev):
    """2^24 items carrying their own push position: every PE's span is
    increasing and names only that PE (size-independent properties)."""
    P, n = 64, 1 << 24
    pos = torch.arange(n, dtype=torch.int64, device="cuda")
    pes = torch.empty(n, dtype=torch.int64, device="cuda")
    B.gen_index(pes, P, seed=9)
    c = Conveyor(P, n, 8)
    c.begin()
    assert c.push(pos, pes)
    assert c.advance(True) == OK
    total = 0
    counts = torch.bincount(pes, minlength=P).cpu()
    for p in range(P):
        got = c.pull(p).view(torch.int64).reshape(-1)
        assert got.numel() == int(counts[p])
        assert bool((got[1:] > got[:-1]).all())
        assert bool((pes[got] == p).all())
        total += got.numel()
    assert total == n
    assert c.advance(True) == DONE
    c.free()


def test_bad_arguments(dev):
    with pytest.raises(B.BaleError):
        Conveyor(0, 10, 8)
    with pytest.raises(B.BaleError):
        Conveyor(257, 10, 8)
    with pytest.raises(B.BaleError):
        Conveyor(4, 10, 10)
    c = Conveyor(4, 10, 8)
    with pytest.raises(B.BaleError):
        c.push(torch.zeros(1, 8, dtype=torch.uint8, device="cuda"), torch.zeros(1, dtype=torch.int64, device="cuda"))
    c.begin()
    with pytest.raises(B.BaleError):
        c.begin()
    c.free()
