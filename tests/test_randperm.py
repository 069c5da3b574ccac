"""Random permutation by dart throwing (rand_permp_agi, spmat/spmat_agi.upc:56-107)
and by a Fisher-Yates pass (rand_permp_agi_opt,
apps/randperm_src/alternates/randperm_agi_opt.upc).

CPU: the oracle's drand48 restatement against glibc's lrand48 and the published
generator; the oracle's sequential dart loop against an independent Python
loop; and the parallel statement the GPU computes (a dart hits iff it is the
first to choose its slot; its item is the number of hits before it) against
the sequential oracle.  For the Fisher-Yates model: the oracle against an
independent Python loop, closed forms at N = 2 and 3 from glibc's own draws,
and the list/predecessor statement the
GPU computes against the oracle.  GPU: the library's permutations equal to
the oracle's element for element, through the native API and the reference's
drop-ins.
"""
import numpy as np
import pytest

import bale_old_amd as B

A48, C48, MASK48 = 0x5DEECE66D, 0xB, (1 << 48) - 1
INIT_STATE = 0x1234ABCD330E  # drand48's state when srand48 was never called


def py_lrand48(state, n):
    out = []
    for _ in range(n):
        state = (A48 * state + C48) & MASK48
        out.append(state >> 17)
    return out


def py_randperm_sequential(N, state):
    """The reference loop at one PE, with the generator restated in Python."""
    M = 2 * N
    target = [-1] * M
    i = 0
    while i < N:
        state = (A48 * state + C48) & MASK48
        r = (state >> 17) % M
        if target[r] == -1:
            target[r] = i
            i += 1
    return [t for t in target if t != -1]


def np_randperm_parallel(N, state, orc, darts):
    """What the GPU computes, restated with numpy on the oracle's sequence."""
    M = 2 * N
    slot = orc.lrand48_seq(state, darts) % M
    # first dart of every slot, in dart order
    _, first_idx = np.unique(slot, return_index=True)
    hit = np.zeros(darts, bool)
    hit[first_idx] = True
    item = np.cumsum(hit) - 1
    target = np.full(M, -1, np.int64)
    keep = hit & (item < N)
    target[slot[keep]] = item[keep]
    assert keep.sum() == N
    return target[target != -1]


# ---- CPU: the oracle pinned ------------------------------------------------------
@pytest.mark.parametrize("seed", [101892, 1, 0x7FFFFFFF, -5, 123456789012])
def test_lrand48_restatement_matches_glibc(orc, seed):
    want = orc.glibc_lrand48(seed, 1000)
    got = orc.lrand48_seq(orc.srand48_state(seed), 1000)
    assert np.array_equal(got, want)
    # srand48 keeps the low 32 bits of the seed and appends 0x330E
    assert orc.srand48_state(seed) == ((seed & 0xFFFFFFFF) << 16) | 0x330E
    assert list(got[:20]) == py_lrand48(orc.srand48_state(seed), 20)


def test_lrand48_published_first_values():
    # X1 = (0x5DEECE66D * 0x330E + 0xB) mod 2^48 after srand48(0); lrand48 = X1 >> 17
    x1 = (0x5DEECE66D * 0x330E + 0xB) % (1 << 48)
    assert py_lrand48(0x330E, 1)[0] == x1 >> 17 == 366850414


@pytest.mark.parametrize("N,seed", [(1, 3), (2, 101892), (7, 11), (100, 101892), (1000, 77)])
def test_oracle_matches_python_sequential(orc, N, seed):
    perm, darts = orc.randperm_agi(N, seed)
    want = py_randperm_sequential(N, orc.srand48_state(seed))
    assert perm.tolist() == want
    assert orc.is_perm(perm, N)
    assert darts >= N


@pytest.mark.parametrize("N,seed", [(1, 5), (3, 9), (1000, 101892), (100_000, 42)])
def test_parallel_statement_equals_sequential(orc, N, seed):
    perm, darts = orc.randperm_agi(N, seed)
    # the same number of darts the sequential run threw, and more: darts past
    # the N-th hit place nothing
    for extra in (0, 5, 1000):
        got = np_randperm_parallel(N, orc.srand48_state(seed), orc, darts + extra)
        assert np.array_equal(got, perm), extra


def test_dart_count_within_the_planned_margin(orc):
    # the GPU generates 2N ln 2 + 8 sqrt(N) + 4096 darts: far above any run
    for N, seed in ((10_000, 1), (300_000, 2), (1_000_000, 101892)):
        _, darts = orc.randperm_agi(N, seed)
        assert abs(darts - 1.3862944 * N) < 4 * np.sqrt(N) + 100, (N, darts)


def py_shuffle_sequential(N, state):
    """rand_permp_agi_opt at one PE with the generator restated in Python:
    N draws choose the (only) PE, then j = i + draw % (N - i), swap."""
    d = py_lrand48(state, 2 * N)
    t = list(range(N))
    for i in range(N):
        j = i + d[N + i] % (N - i)
        t[i], t[j] = t[j], t[i]
    return t


def np_shuffle_parallel(N, state, orc):
    """What the GPU computes for the Fisher-Yates model: per-target lists of
    steps, prev(i) = the last earlier step with i's target, pred(k) = the last
    earlier step targeting k, perm[i] = chase(prev(i)) or j_i."""
    r = orc.lrand48_seq(state, 2 * N)[N:]
    i = np.arange(N)
    j = i + r % (N - i)
    order = np.lexsort((i, j))  # by target, then step
    js, ks = j[order], i[order]
    same = np.r_[False, js[1:] == js[:-1]]
    prev = np.full(N, -1)
    prev[ks[same]] = ks[np.nonzero(same)[0] - 1]
    pred = np.full(N, -1)
    for p_, k_ in zip(js, ks):  # the last member below p of p's list
        if k_ < p_:
            pred[p_] = max(pred[p_], k_)
    perm = j.copy()
    for a in np.nonzero(prev >= 0)[0]:
        q = prev[a]
        while pred[q] >= 0:
            q = pred[q]
        perm[a] = q
    return perm


@pytest.mark.parametrize("N,seed", [(1, 3), (2, 101892), (7, 11), (100, 101892), (3000, -7)])
def test_shuffle_oracle_matches_python_sequential(orc, N, seed):
    perm = orc.randperm_agi_opt(N, seed)
    assert perm.tolist() == py_shuffle_sequential(N, orc.srand48_state(seed))
    assert orc.is_perm(perm, N)


@pytest.mark.parametrize("seed", [1, 2, 3, 101892, 99])
def test_shuffle_oracle_closed_form_at_two(orc, seed):
    """N = 2: draws 0 and 1 choose the PE; draw 2 swaps the items iff odd;
    draw 3 picks j_1 = 1.  The draws are glibc's own."""
    d = orc.glibc_lrand48(seed, 4)
    want = [1, 0] if d[2] % 2 else [0, 1]
    assert orc.randperm_agi_opt(2, seed).tolist() == want


# N = 3, worked by hand from 0 1 2: the swaps (0, j0) then (1, j1) give
FY3 = {(0, 1): [0, 1, 2], (0, 2): [0, 2, 1], (1, 1): [1, 0, 2],
       (1, 2): [1, 2, 0], (2, 1): [2, 1, 0], (2, 2): [2, 0, 1]}


def test_shuffle_oracle_known_answers_at_three(orc):
    """N = 3 over many seeds: draws 0-2 choose the PE, j0 = draw 3 mod 3,
    j1 = 1 + draw 4 mod 2 (glibc's own draws); the order is the hand-worked
    table's, and every row of the table is met.  (The orders are not equally
    likely: small moduli read the generator's low bits, which successive
    draws share, in the reference as here.)"""
    seen = set()
    for seed in range(1, 400):
        d = orc.glibc_lrand48(seed, 6)
        key = (int(d[3] % 3), int(1 + d[4] % 2))
        assert orc.randperm_agi_opt(3, seed).tolist() == FY3[key], seed
        seen.add(key)
    assert seen == set(FY3)


@pytest.mark.parametrize("N,seed", [(1, 5), (2, 8), (3, 9), (1000, 101892), (50_000, 42)])
def test_shuffle_parallel_statement_equals_sequential(orc, N, seed):
    want = orc.randperm_agi_opt(N, seed)
    assert np.array_equal(np_shuffle_parallel(N, orc.srand48_state(seed), orc), want)


def test_oracle_is_perm_detects_errors(orc):
    p = np.arange(10, dtype=np.int64)
    assert orc.is_perm(p, 10)
    q = p.copy()
    q[3] = 4
    assert not orc.is_perm(q, 10)
    q = p.copy()
    q[9] = 10
    assert not orc.is_perm(q, 10)


# ---- GPU -----------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("N", [1, 2, 3, 1000, 12_345, 1 << 20])
@pytest.mark.parametrize("seed", [101892, 7, -3])
def test_randperm_matches_oracle(dev, orc, N, seed):
    import torch
    perm = torch.full((N,), -9, dtype=torch.int64, device=dev)
    B.randperm(perm, seed)
    want, _ = orc.randperm_agi(N, seed)
    assert np.array_equal(perm.cpu().numpy(), want)
    assert B.is_perm(perm)


@pytest.mark.gpu
@pytest.mark.parametrize("N,seed", [(1 << 26, 101892), ((1 << 26) - 12345, 5), ((1 << 22) - 1, 3),
                                    ((1 << 22) + 7, -3)])
def test_randperm_by_region_matches_oracle(dev, orc, N, seed):
    """Permutations resolved by slot region, which runs from one 2^15-slot
    region per CU (2^22 - 1 elements: 256 regions, the last one short) to
    2^27 slots (2^26 elements: 4096 regions, ~9.3e7 darts in 5679 chunks),
    element for element."""
    import torch
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm(perm, seed)
    want, _ = orc.randperm_agi(N, seed)
    assert np.array_equal(perm.cpu().numpy(), want)


@pytest.mark.gpu
def test_randperm_large_matches_oracle(dev, orc):
    """2^27 elements (2^28 slots, ~1.9e8 darts: first darts by device atomic
    min), element for element."""
    import torch
    N = 1 << 27
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm(perm, 101892)
    want, _ = orc.randperm_agi(N, 101892)
    assert np.array_equal(perm.cpu().numpy(), want)


@pytest.mark.gpu
def test_randperm_near_the_size_limit(dev):
    """2^29 elements (2^30 slots, 7.4e8 darts): a permutation, and the same
    one on a second call."""
    import torch
    N = 1 << 29
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm(perm, 7)
    assert B.is_perm(perm)
    head, tail = perm[:4096].clone(), perm[-4096:].clone()
    B.randperm(perm, 7)
    assert bool((perm[:4096] == head).all()) and bool((perm[-4096:] == tail).all())
    del perm
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_randperm_seed0_uses_the_initial_state(dev, orc):
    import torch
    N = 5000
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm(perm, 0)
    want = np.array(py_randperm_sequential(N, INIT_STATE), dtype=np.int64)
    assert np.array_equal(perm.cpu().numpy(), want)


@pytest.mark.gpu
def test_randperm_dropins_and_is_perm(dev, orc):
    L = B.lib()
    N, seed = 100_000, 101892  # the reference driver's seed at PE 0 (randperm.upc:107)
    want, _ = orc.randperm_agi(N, seed)
    for name, extra in (("rand_permp_agi", ()), ("rand_permp_exstack", (1024,)),
                        ("rand_permp_exstack2", (1024,)), ("rand_permp_conveyor", ())):
        p = getattr(L, name)(N, seed, *extra)
        assert p, B.last_error()
        gpu_s, _, _ = B.last_timing()
        assert gpu_s > 0
        host = np.empty(N, np.int64)
        assert L.bale_memget(host.ctypes.data, p, N * 8) == 0
        assert np.array_equal(host, want), name
        assert L.is_perm(p, N) == 1
        L.bale_all_free(p)
    assert L.rand_permp_agi(-1, 1) is None
    assert "outside" in B.last_error()
    # is_perm on host arrays (staged to the GPU)
    assert L.is_perm(want.ctypes.data, N) == 1
    bad = want.copy()
    bad[17] = bad[18]
    assert L.is_perm(bad.ctypes.data, N) == 0


@pytest.mark.gpu
def test_is_perm_on_gpu_detects_errors(dev):
    import torch
    p = torch.arange(1000, dtype=torch.int64, device=dev)
    assert B.is_perm(p)
    q = p.clone()
    q[10] = 11
    assert not B.is_perm(q)
    q = p.clone()
    q[5] = -1
    assert not B.is_perm(q)
    q = p.clone()
    q[0] = 1000
    assert not B.is_perm(q)
    assert B.is_perm(torch.empty(0, dtype=torch.int64, device=dev))


# ---- GPU: the Fisher-Yates model (rand_permp_agi_opt) ---------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("N", [1, 2, 3, 1000, 12_345, 1 << 20])
@pytest.mark.parametrize("seed", [101892, 7, -3])
def test_shuffle_matches_oracle(dev, orc, N, seed):
    import torch
    perm = torch.full((N,), -9, dtype=torch.int64, device=dev)
    B.randperm_shuffle(perm, seed)
    assert np.array_equal(perm.cpu().numpy(), orc.randperm_agi_opt(N, seed))


@pytest.mark.gpu
def test_shuffle_large_matches_oracle(dev, orc):
    """2^26 elements, element for element."""
    import torch
    N = 1 << 26
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm_shuffle(perm, 101892)
    assert np.array_equal(perm.cpu().numpy(), orc.randperm_agi_opt(N, 101892))


@pytest.mark.gpu
def test_shuffle_near_the_size_limit(dev):
    """2^30 - 1 elements: a permutation, the same one on a second call."""
    import torch
    N = (1 << 30) - 1
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm_shuffle(perm, 5)
    assert B.is_perm(perm)
    head, tail = perm[:4096].clone(), perm[-4096:].clone()
    B.randperm_shuffle(perm, 5)
    assert bool((perm[:4096] == head).all()) and bool((perm[-4096:] == tail).all())
    del perm
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_shuffle_seed0_and_dropin(dev, orc):
    import torch
    N = 4000
    perm = torch.empty(N, dtype=torch.int64, device=dev)
    B.randperm_shuffle(perm, 0)
    assert perm.cpu().tolist() == py_shuffle_sequential(N, INIT_STATE)
    L = B.lib()
    N, seed = 100_000, 101892
    p = L.rand_permp_agi_opt(N, seed)
    assert p, B.last_error()
    assert B.last_timing()[0] > 0
    host = np.empty(N, np.int64)
    assert L.bale_memget(host.ctypes.data, p, N * 8) == 0
    assert np.array_equal(host, orc.randperm_agi_opt(N, seed))
    L.bale_all_free(p)
    assert L.rand_permp_agi_opt(1 << 30, 1) is None
    assert L.bale_randperm_shuffle(None, 1 << 30, 1, None) == B.ERR_ARG
    assert L.bale_randperm_shuffle(None, -1, 1, None) == B.ERR_ARG
