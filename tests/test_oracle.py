"""Pins for the CPU oracle (oracle/bale_oracle.c).

The reference ships no golden vectors (its only pins are self-checks), so each
oracle function is pinned here by something independent of its own code:
published known-answer values, Python big-integer arithmetic, numpy's textbook
routines, exact probability mass functions, brute force on small inputs, and
the reference drivers' own closed-form and zero-sum checks.
"""
import math

import numpy as np
import pytest

M64 = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def mix_py(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


# ---- splitmix64 --------------------------------------------------------------
# Published first outputs of splitmix64 (Vigna's reference implementation).
KAT = {
    0: [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F],
    1: [0x910A2DEC89025CC1, 0xBEEB8DA1658EEC67, 0xF893A2EEFB32555E],
}


def test_splitmix64_known_answers(orc):
    import ctypes as C
    for seed, outs in KAT.items():
        st = C.c_uint64(seed)
        got = [orc.lib().or_splitmix64_next(C.byref(st)) for _ in outs]
        assert got == outs
        assert [orc.splitmix64_at(seed, i) for i in range(3)] == outs
    assert orc.splitmix64(0) == KAT[0][0]
    assert orc.splitmix64(1) == KAT[1][0]


def test_reduce_matches_bigint_multiply_high(orc):
    rng = np.random.default_rng(7)
    xs = [int(v) for v in rng.integers(0, 2**63, 200, dtype=np.uint64)] + [0, M64, 1 << 63]
    for n in [1, 2, 3, 7, 1000, 1 << 20, (1 << 20) + 1, 1 << 28, 12345678901]:
        for x in xs:
            assert orc.reduce(x, n) == (x * n) >> 64, (x, n)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 17, 100, 1000, 4096, 5000])
def test_perm_is_a_bijection(orc, n):
    out = [orc.perm(x, n, 12345) for x in range(n)]
    assert sorted(out) == list(range(n))
    if n >= 100:
        assert sum(1 for x, y in zip(range(n), out) if x == y) < n // 10
        # a different key gives a different permutation
        assert out != [orc.perm(x, n, 999) for x in range(n)]


def test_gen_uniform_matches_stream(orc):
    for n, rng, seed in [(1000, 1 << 20, 1), (1000, 1000, 5), (257, 1 << 28, 2), (100, 1, 3)]:
        got = orc.gen_uniform(n, rng, seed)
        want = [((mix_py((seed + (i + 1) * G) & M64)) * rng) >> 64 for i in range(n)]
        assert got.tolist() == want
        # the general generator with dist 0 and no runs is the same stream
        assert np.array_equal(orc.gen_index(n, rng, 0, 1.0, 0, seed), got)


def chi2_pvalue_ok(obs, exp_p, n):
    exp = exp_p * n
    keep = exp >= 5
    o = np.concatenate([obs[keep], [obs[~keep].sum()]])
    e = np.concatenate([exp[keep], [exp[~keep].sum()]])
    if e[-1] == 0:
        o, e = o[:-1], e[:-1]
    stat = ((o - e) ** 2 / e).sum()
    from scipy.stats import chi2
    return chi2.sf(stat, len(o) - 1)


@pytest.mark.parametrize("nbins,s", [(50, 1.1), (50, 0.8), (200, 1.4), (30, 1.0), (1000, 1.1)])
def test_zipf_matches_exact_pmf(orc, nbins, s):
    n = 200_000
    seed = 11
    vals = orc.gen_index(n, nbins, 1, s, 0, seed)
    assert vals.min() >= 0 and vals.max() < nbins
    # undo the keyed permutation to recover ranks
    pkey = mix_py(seed ^ 0x5045524D5045524D)
    inv = np.empty(nbins, np.int64)
    for r in range(nbins):
        inv[orc.perm(r, nbins, pkey)] = r + 1
    ranks = inv[vals]
    obs = np.bincount(ranks, minlength=nbins + 1)[1:].astype(float)
    pmf = np.array([k ** -s for k in range(1, nbins + 1)])
    pmf /= pmf.sum()
    assert chi2_pvalue_ok(obs, pmf, n) > 1e-4
    # the oracle's own pmf agrees with the closed form
    assert math.isclose(orc.zipf_pmf(1, nbins, s), pmf[0], rel_tol=1e-12)
    # a sign or off-by-one mistake in the sampler moves the top rank's mass a lot
    assert abs(obs[0] / n - pmf[0]) < 5 * math.sqrt(pmf[0] * (1 - pmf[0]) / n) + 1e-3


def test_zipf_large_table_top_mass(orc):
    # 2^26 bins, s = 1.1: the hottest bin holds 1/H(2^26, 1.1) of the mass
    nbins, s, n = 1 << 26, 1.1, 400_000
    vals = orc.gen_index(n, nbins, 1, s, 0, 3)
    top = np.bincount(vals % (1 << 20)).max()  # hottest bin dominates its residue class
    k = np.arange(1, 1 << 22, dtype=np.float64)
    H = (k ** -s).sum() + ((1 << 22) ** (1 - s) - nbins ** (1 - s)) / (s - 1)
    assert abs(top / n - 1 / H) < 0.01


def test_clustered_runs(orc):
    n, nbins, mean = 200_000, 1 << 20, 64
    v = orc.gen_index(n, nbins, 0, 1.0, mean, 4)
    starts = np.flatnonzero(np.r_[True, v[1:] != v[:-1]])
    lengths = np.diff(np.r_[starts, n])
    # runs are cut at multiples of 4096 as well, so the mean is a little lower
    assert 50 < lengths.mean() < 70
    assert lengths.max() <= 4096
    # every run's value is the uniform draw at the run's first element
    plain = orc.gen_index(n, nbins, 0, 1.0, 0, 4)
    assert np.array_equal(v[starts], plain[starts])


# ---- reference inputs --------------------------------------------------------
def glibc_rand(seed, count):
    """glibc random(3), TYPE_3 additive feedback generator (published algorithm)."""
    r = [0] * (344 + count)
    r[0] = seed if seed != 0 else 1
    for i in range(1, 31):
        hi, lo = divmod(r[i - 1], 127773)
        word = 16807 * lo - 2836 * hi
        if word < 0:
            word += 2147483647
        r[i] = word
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + count):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return [r[i] >> 1 for i in range(344, 344 + count)]


def test_reference_rand_inputs(orc):
    got = orc.reference_rand(5000, 1000, 120348)  # histo -T 1000 at one PE (histo.upc:143-146)
    want = [x % 1000 for x in glibc_rand(120348, 5000)]
    assert got.tolist() == want
    got = orc.reference_rand(3000, 100000, 5)  # ig default table (ig.upc:168-170)
    assert got.tolist() == [x % 100000 for x in glibc_rand(5, 3000)]


def test_pack_closed_form(orc):
    idx = np.arange(0, 100_000, 37, dtype=np.int64)
    for P in (1, 3, 16):
        got = orc.pack(idx, P)
        assert np.array_equal(got, ((idx // P) << 16) | (idx % P))


# ---- histogram ---------------------------------------------------------------
@pytest.mark.parametrize("nbins,n", [(1, 10), (7, 1000), (1000, 100_000), (1 << 16, 1 << 18)])
def test_histo_agi_matches_bincount(orc, nbins, n):
    idx = orc.gen_uniform(n, nbins, 9)
    base = np.random.default_rng(1).integers(0, 100, nbins).astype(np.int64)
    counts = base.copy()
    orc.histo_agi(idx, counts)
    assert np.array_equal(counts, base + np.bincount(idx, minlength=nbins))
    c32 = base.astype(np.int32)
    orc.histo_agi(idx, c32)
    assert np.array_equal(c32, counts.astype(np.int32))


@pytest.mark.parametrize("P", [1, 3, 8])
def test_histo_owner_matches_agi_any_pe_count(orc, P):
    nbins, n = 10_000, 200_000
    idx = orc.gen_index(n, nbins, 1, 1.1, 16, 21)
    lnum = (nbins + P - 1) // P
    lcounts = np.zeros(lnum * P, np.int64)
    assert orc.histo_owner(orc.pack(idx, P), lcounts, P) == 0
    glob = np.zeros(lnum * P, np.int64)
    for pe in range(P):
        glob[np.arange(lnum) * P + pe] = lcounts[pe * lnum:(pe + 1) * lnum]
    assert np.array_equal(glob[:nbins], np.bincount(idx, minlength=nbins))


def test_histo_owner_rejects_foreign_pe(orc):
    pck = np.array([(5 << 16) | 2], np.int64)
    assert orc.histo_owner(pck, np.zeros(16, np.int64), 2) == -1


def test_reference_zero_sum_check(orc):
    # histo.upc:208-227: every model adds 1 per index into the same counts;
    # subtracting num_models per index must leave zeros
    nbins, n = 5000, 50_000
    idx = orc.reference_rand(n, nbins, 120348)
    counts = np.zeros(nbins, np.int64)
    orc.histo_agi(idx, counts)
    orc.histo_owner(orc.pack(idx, 1), counts, 1)
    np.subtract.at(counts, idx, 2)
    assert not counts.any()


def test_histo_mt_matches(orc):
    idx = orc.gen_uniform(300_000, 4096, 3)
    counts = np.zeros(4096, np.int64)
    secs = orc.histo_agi_mt(idx, counts, 4)
    assert secs >= 0
    assert np.array_equal(counts, np.bincount(idx, minlength=4096))


@pytest.mark.parametrize("P", [1, 2, 5, 8])
def test_histo_owner_mt_matches(orc, P):
    """The threaded aggregate-by-owner baseline (bulk-synchronous rounds of at
    most 1024 items per destination) gives numpy's bincount, including skewed
    input whose hot owner fills a buffer every round."""
    for idx in (orc.gen_uniform(300_000, 4099, 3), orc.gen_index(200_000, 5000, 1, 1.4, 0, 4),
                np.zeros(5000, np.int64), np.zeros(0, np.int64)):
        counts = np.full(5000, 2, np.int64)
        assert orc.histo_owner_mt(idx, counts, P) >= 0
        want = np.bincount(idx, minlength=5000) + 2
        assert np.array_equal(counts, want), P


# ---- index gather ------------------------------------------------------------
def test_fill_closed_forms(orc):
    t = orc.fill_splitmix(1000)
    assert [int(x) & M64 for x in t] == [mix_py((j + G) & M64) for j in range(1000)]
    assert np.array_equal(orc.fill_neg(10), -(np.arange(10) + 1))


def test_ig_agi_matches_numpy_and_closed_forms(orc):
    tab_len, n = 100_000, 300_000
    idx = orc.reference_rand(n, tab_len, 5)
    neg = orc.fill_neg(tab_len)
    tgt = orc.ig_agi(idx, neg)
    assert np.array_equal(tgt, -(idx + 1))  # ig.upc:96-114
    sm = orc.fill_splitmix(tab_len)
    assert np.array_equal(orc.ig_agi(idx, sm), sm[idx])


@pytest.mark.parametrize("P", [1, 4, 7])
def test_ig_request_reply_models(orc, P):
    tab_len, n = 9973, 50_000
    table = orc.fill_splitmix(tab_len)
    idx = orc.gen_index(n, tab_len, 1, 1.2, 8, 6)
    lnum = (tab_len + P - 1) // P
    ltables = np.zeros(lnum * P, np.int64)
    for pe in range(P):
        js = np.arange(lnum) * P + pe
        ok = js < tab_len
        ltables[pe * lnum:(pe + 1) * lnum][ok] = table[js[ok]]
    pck = orc.pack(idx, P)
    want = table[idx]
    assert np.array_equal(orc.ig_exstack(pck, ltables, P), want)
    assert np.array_equal(orc.ig_conveyor(pck, ltables, P), want)


def test_ig_mt_matches(orc):
    table = orc.fill_splitmix(50_000)
    idx = orc.gen_uniform(200_000, 50_000, 8)
    tgt = np.zeros_like(idx)
    orc.ig_agi_mt(idx, table, tgt, 4)
    assert np.array_equal(tgt, table[idx])


@pytest.mark.parametrize("P", [1, 3, 8])
def test_ig_owner_mt_matches(orc, P):
    """The threaded request/reply baseline (requests to owners, replies in
    place, retraced in order per round) gives table[index]."""
    table = orc.fill_splitmix(50_001)
    for idx in (orc.gen_uniform(200_000, 50_001, 8), orc.gen_index(120_000, 50_001, 1, 1.3, 16, 9)):
        tgt = np.full_like(idx, -1)
        assert orc.ig_owner_mt(idx, table, tgt, P) >= 0
        assert np.array_equal(tgt, table[idx]), P


# ---- batched mix parameters -----------------------------------------------------
def test_batch_params_ranges(orc):
    inst = orc.batch_params(4096, 4)
    assert inst == orc.batch_params(4096, 4)
    tl = np.array([x["table_len"] for x in inst])
    no = np.array([x["nops"] for x in inst])
    assert tl.min() >= 1 << 10 and tl.max() < 1 << 24
    assert no.min() >= 1 << 12 and no.max() < 1 << 22
    # log-uniform: mean log2 near the middle of the range
    assert abs(np.log2(tl).mean() - 17) < 0.3
    assert abs(np.log2(no).mean() - 17) < 0.3
    kinds = np.array([x["kind"] for x in inst])
    assert abs(kinds.mean() - 0.5) < 0.05
    dists = np.bincount([x["dist"] for x in inst], minlength=3) / 4096
    assert np.all(np.abs(dists - 1 / 3) < 0.05)
    s = np.array([x["s"] for x in inst])
    assert s.min() >= 0.8 and s.max() < 1.4


def test_digest_is_position_sensitive(orc):
    a = np.arange(100, dtype=np.int64)
    b = a.copy()
    b[[3, 4]] = b[[4, 3]]
    assert orc.digest(a) != orc.digest(b)
    assert orc.digest(a) == orc.digest(a.copy())


# ---- conveyor delivery ---------------------------------------------------------------
def _fifo_brute(items, pes, P):
    """Per-destination FIFO queues, filled in push order (textbook restatement)."""
    q = [[] for _ in range(P)]
    dropped = 0
    for row, p in zip(items, pes):
        if 0 <= p < P:
            q[p].append(bytes(row))
        else:
            dropped += 1
    return q, dropped


@pytest.mark.parametrize("P,item_bytes", [(1, 8), (7, 16), (64, 4), (256, 32)])
def test_convey_deliver_is_stable_partition(orc, P, item_bytes):
    rng = np.random.default_rng(P * 100 + item_bytes)
    n = 20_000
    items = rng.integers(0, 256, size=(n, item_bytes), dtype=np.uint8)
    pes = rng.integers(-2, P + 2, size=n).astype(np.int64)
    out, starts, dropped = orc.convey_deliver(items, pes, P)
    ok = (pes >= 0) & (pes < P)
    assert dropped == int((~ok).sum())
    # library routine: a stable sort of the valid items by destination
    order = np.argsort(pes[ok], kind="stable")
    assert np.array_equal(out, items[ok][order])
    assert np.array_equal(starts, np.concatenate([[0], np.cumsum(np.bincount(pes[ok], minlength=P))]))


def test_convey_deliver_brute_force(orc):
    rng = np.random.default_rng(3)
    for trial in range(20):
        P = int(rng.integers(1, 9))
        n = int(rng.integers(0, 40))
        items = rng.integers(0, 256, size=(max(n, 1), 8), dtype=np.uint8)[:n]
        pes = rng.integers(-1, P + 1, size=n).astype(np.int64)
        if n == 0:
            continue
        out, starts, dropped = orc.convey_deliver(items, pes, P)
        q, d2 = _fifo_brute(items, pes, P)
        assert dropped == d2
        for p in range(P):
            got = [bytes(r) for r in out[starts[p]:starts[p + 1]]]
            assert got == q[p], (trial, p)
