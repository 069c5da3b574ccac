"""Sparse matrices: Erdos-Renyi generation, transpose (transpose_matrix_agi,
spmat/spmat_agi.upc:176-236) and permutation (permute_matrix_agi,
spmat/spmat_agi.upc:119-169).

CPU: the oracle pinned -- its log restatement against libm, the generator's
structure and density, its transpose against scipy's CSR transpose and
brute force, its permutation against a dense restatement and a row-by-row
Python restatement.  GPU: the library equal to the oracle on the same
matrices, element for element, through the native API and the drop-ins.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import bale_old_amd as B


def make_csr(rows):
    off = np.zeros(len(rows) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nz = np.array([c for r in rows for c in r], np.int64)
    return off, nz


def rand_rows(rng, numrows, numcols, mean_len, dup=False):
    rows = []
    for _ in range(numrows):
        k = rng.poisson(mean_len)
        if dup:
            rows.append(list(rng.integers(0, numcols, k)))
        else:
            k = min(k, numcols)
            rows.append(list(rng.choice(numcols, k, replace=False)))
    return make_csr(rows)


def py_transpose(numrows, numcols, off, nz):
    """The reference's loop (rows in order, next free slot per column) with lists."""
    out = [[] for _ in range(numcols)]
    for r in range(numrows):
        for c in nz[off[r]:off[r + 1]]:
            out[c].append(r)
    return make_csr(out)


def py_permute(off, nz, rperminv, cperminv):
    n = off.size - 1
    out = [None] * n
    for i in range(n):
        out[rperminv[i]] = [cperminv[c] for c in nz[off[i]:off[i + 1]]]
    return make_csr(out)


# ---- CPU: the oracle pinned ------------------------------------------------------
def test_det_log_matches_libm(orc):
    rng = np.random.default_rng(7)
    xs = [1.0, 0.5, 0.25, 2.0 ** -53, 1 - 2.0 ** -53, 0.7071067811865476, 0.7071067811865475]
    xs += list(rng.random(20000) + 1e-300)
    for x in xs:
        a, b = orc.det_log(float(x)), math.log(float(x))
        assert abs(a - b) <= 4e-16 * max(1.0, abs(b)), (x, a, b)
    assert orc.det_log(1.0) == 0.0


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("unit_diag", [0, 1])
def test_gen_er_structure(orc, mode, unit_diag):
    n, p = 3000, 0.004
    off, nz = orc.gen_er(n, p, unit_diag, mode, 11)
    assert off[0] == 0 and np.all(np.diff(off) >= 0) and off[-1] == nz.size
    nlo = nhi = 0
    for r in range(n):
        row = nz[off[r]:off[r + 1]]
        assert np.all(np.diff(row) > 0)  # strictly ascending
        assert np.all((row >= 0) & (row < n))
        assert (r in row) == bool(unit_diag)
        nlo += np.sum(row < r)
        nhi += np.sum(row > r)
        if mode == 1:
            assert np.all(row <= r)
        if mode == 2:
            assert np.all(row >= r)
    pairs = n * (n - 1) / 2
    for got, on in ((nlo, mode in (1, 3)), (nhi, mode in (2, 3))):
        if on:  # binomial(pairs, p): within 5 standard deviations
            assert abs(got - pairs * p) < 5 * math.sqrt(pairs * p), (got, pairs * p)
        else:
            assert got == 0


def test_gen_er_halves_and_extremes(orc):
    n, p, seed = 500, 0.03, 5
    lo, lnz = orc.gen_er(n, p, 0, 1, seed)
    up, unz = orc.gen_er(n, p, 1, 2, seed)
    both, bnz = orc.gen_er(n, p, 1, 3, seed)
    for r in range(n):  # mode 3 = lower half of mode 1, diagonal, upper half of mode 2
        want = list(lnz[lo[r]:lo[r + 1]]) + list(unz[up[r]:up[r + 1]])
        assert list(bnz[both[r]:both[r + 1]]) == want
    off, nz = orc.gen_er(40, 1.0, 0, 3, 1)  # p = 1: every off-diagonal entry
    assert off[-1] == 40 * 39
    off, nz = orc.gen_er(40, 0.0, 1, 3, 1)  # p = 0: the diagonal only
    assert list(nz) == list(range(40))


def test_gen_er_gaps_are_geometric(orc):
    # gaps between consecutive upper-half entries: P(gap = k) = p (1 - p)^k
    # (a gap is only seen whole when the row has room for it: gaps starting
    # at least L columns before the row's end are counted, those of L or more
    # in one bin)
    n, p, L = 4000, 0.01, 600
    off, nz = orc.gen_er(n, p, 0, 2, 3)
    gaps = []
    for r in range(n):
        s = r + 1
        for c in nz[off[r]:off[r + 1]]:
            if n - s > L:
                gaps.append(min(c - s, L))
            s = c + 1
        if n - s > L:
            gaps.append(L)
    gaps = np.array(gaps)
    edges = [0, 10, 30, 60, 100, 200, 400, L, L + 1]
    obs = np.histogram(gaps, edges)[0]
    q = 1 - p
    prob = [q ** a - q ** b for a, b in zip(edges[:-2], edges[1:-1])] + [q ** L]
    exp = np.array(prob) * gaps.size
    chi2 = float(np.sum((obs - exp) ** 2 / exp))
    assert chi2 < 30, (obs, exp)  # 7 degrees of freedom


def py_det_log(x):
    """The oracle's logarithm restated in Python floats, operation for
    operation (Python floats are IEEE doubles and Python does not fuse)."""
    import struct
    b = struct.unpack("<Q", struct.pack("<d", x))[0]
    e = ((b >> 52) & 0x7FF) - 1022
    m = struct.unpack("<d", struct.pack("<Q", (b & 0x000FFFFFFFFFFFFF) | (0x3FE << 52)))[0]
    if m < 0.70710678118654752440:
        m = m * 2.0
        e -= 1
    s = (m - 1.0) / (m + 1.0)
    z = s * s
    t = 1.0 / 25.0
    for k in range(11, -1, -1):
        t = t * z + 1.0 / float(2 * k + 1)
    return float(e) * 0.69314718055994530942 + 2.0 * s * t


def py_gen_er(n, p, unit_diag, mode, seed):
    """The generator restated in Python: per (row, half) a splitmix64 stream,
    geometric gaps floor(log(u) / log(1 - p)), u uniform in (0, 1]."""
    M = (1 << 64) - 1
    G = 0x9E3779B97F4A7C15

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    D = py_det_log(1.0 - p)

    def half(row, h, lo, hi):
        key = mix((seed + G * (2 * row + h + 1) + G) & M)
        out, col, i = [], lo - 1, 0
        while True:
            x = mix((key + (i + 1) * G) & M)
            i += 1
            u = float((x >> 11) + 1) * 2.0 ** -53
            g = math.floor(py_det_log(u) / D)
            if g >= hi - col:
                return out
            col += 1 + g
            if col >= hi:
                return out
            out.append(col)

    rows = []
    for r in range(n):
        row = half(r, 0, 0, r) if mode in (1, 3) else []
        if unit_diag:
            row.append(r)
        if mode in (2, 3):
            row += half(r, 1, r + 1, n)
        rows.append(row)
    return make_csr(rows)


@pytest.mark.parametrize("mode,unit_diag", [(1, 0), (2, 1), (3, 0), (3, 1)])
def test_gen_er_matches_python_restatement(orc, mode, unit_diag):
    n, p, seed = 300, 0.05, 1234567
    off, nz = orc.gen_er(n, p, unit_diag, mode, seed)
    woff, wnz = py_gen_er(n, p, unit_diag, mode, seed)
    assert np.array_equal(off, woff) and np.array_equal(nz, wnz)
    rng = np.random.default_rng(5)
    for x in rng.random(2000) + 1e-12:
        assert orc.det_log(float(x)) == py_det_log(float(x))


def test_gen_er_gap_probabilities_exact(orc):
    # p = 1/2: a gap is 0 with probability 1/2 and 1 with probability 1/4 (a
    # shifted gap or a wrong logarithm base moves these by far more than the
    # 5-sigma band)
    n = 1500
    off, nz = orc.gen_er(n, 0.5, 0, 2, 9)
    g0 = g1 = tot = 0
    for r in range(n - 40):
        row = nz[off[r]:off[r + 1]]
        prev = r
        for c in row:
            if c >= n - 20:
                break
            g = c - prev - 1
            g0 += g == 0
            g1 += g == 1
            tot += 1
            prev = c
    for got, pr in ((g0, 0.5), (g1, 0.25)):
        assert abs(got - tot * pr) < 5 * math.sqrt(tot * pr * (1 - pr)), (got, tot * pr)


@pytest.mark.parametrize("numrows,numcols,mean,dup", [
    (0, 5, 1, False), (6, 0, 0, False), (1, 1, 1, False), (50, 50, 4, False),
    (200, 37, 6, True), (300, 1000, 3, False), (17, 300, 40, True)])
def test_oracle_transpose(orc, numrows, numcols, mean, dup):
    rng = np.random.default_rng(numrows * 7 + numcols)
    off, nz = rand_rows(rng, numrows, max(numcols, 1), mean if numcols else 0, dup)
    toff, tnz = orc.transpose_agi(numrows, numcols, off, nz)
    woff, wnz = py_transpose(numrows, numcols, off, nz)
    assert np.array_equal(toff, woff) and np.array_equal(tnz, wnz)
    if numrows and numcols and not dup:
        A = sp.csr_matrix((np.ones(nz.size), nz, off), shape=(numrows, numcols))
        T = A.T.tocsr()
        T.sort_indices()
        assert np.array_equal(T.indptr, toff) and np.array_equal(T.indices, tnz)


def test_oracle_transpose_twice_sorts_rows(orc):
    rng = np.random.default_rng(3)
    off, nz = rand_rows(rng, 400, 300, 5)
    toff, tnz = orc.transpose_agi(400, 300, off, nz)
    ooff, onz = orc.transpose_agi(300, 400, toff, tnz)
    assert np.array_equal(ooff, off)
    for r in range(400):
        assert list(onz[ooff[r]:ooff[r + 1]]) == sorted(nz[off[r]:off[r + 1]])


@pytest.mark.parametrize("n,mean", [(1, 1), (30, 3), (500, 6)])
def test_oracle_permute(orc, n, mean):
    rng = np.random.default_rng(n)
    off, nz = rand_rows(rng, n, n, mean)
    rpi, cpi = rng.permutation(n).astype(np.int64), rng.permutation(n).astype(np.int64)
    poff, pnz = orc.permute_agi(off, nz, rpi, cpi)
    woff, wnz = py_permute(off, nz, rpi, cpi)
    assert np.array_equal(poff, woff) and np.array_equal(pnz, wnz)
    A = sp.csr_matrix((np.ones(nz.size), nz, off), shape=(n, n)).toarray()
    Pm = sp.csr_matrix((np.ones(pnz.size), pnz, poff), shape=(n, n)).toarray()
    want = np.zeros_like(A)
    want[np.ix_(rpi, cpi)] = A  # Ap[rperminv[i], cperminv[j]] = A[i, j]
    assert np.array_equal(Pm, want)


# ---- GPU -----------------------------------------------------------------------------------
def _dev(t):
    import torch
    return torch.from_numpy(np.ascontiguousarray(t)).to("cuda:0")


@pytest.mark.gpu
@pytest.mark.parametrize("n,p,unit_diag,mode", [
    (1, 0.5, 1, 3), (100, 0.0, 1, 3), (60, 1.0, 0, 3), (2000, 0.01, 0, 1), (2000, 0.01, 1, 2),
    (50000, 18 / 50000, 0, 3), (1 << 18, 18 / (1 << 18), 1, 2)])
def test_gen_er_matches_oracle(dev, orc, n, p, unit_diag, mode):
    off, nz = B.spmat_gen_er(n, p, unit_diag, mode, seed=101894)
    woff, wnz = orc.gen_er(n, p, unit_diag, mode, 101894)
    assert np.array_equal(off.cpu().numpy(), woff)
    assert np.array_equal(nz.cpu().numpy(), wnz)


TRANSPOSE_CASES = [
    # (numrows, numcols, mean row length, repeats in a row)
    (0, 7, 0, False), (5, 0, 0, False), (1, 1, 1, False), (3, 2, 2, True),
    (1000, 255, 4, False), (1000, 256, 4, False), (1000, 257, 4, True),
    (20000, 70000, 7, False), (5000, 3, 30, True), (300, 1 << 20, 50, False),
]


@pytest.mark.gpu
@pytest.mark.parametrize("numrows,numcols,mean,dup", TRANSPOSE_CASES)
def test_transpose_matches_oracle(dev, orc, numrows, numcols, mean, dup):
    rng = np.random.default_rng(numrows + 3 * numcols)
    if numrows * mean > 100000:
        lens = rng.poisson(mean, numrows)
        off = np.zeros(numrows + 1, np.int64)
        off[1:] = np.cumsum(lens)
        nz = rng.integers(0, numcols, off[-1], dtype=np.int64)  # repeats possible, any order
    else:
        off, nz = rand_rows(rng, numrows, max(numcols, 1), mean if numcols else 0, dup)
    toff, tnz = B.spmat_transpose(numcols, _dev(off), _dev(nz))
    woff, wnz = orc.transpose_agi(numrows, numcols, off, nz)
    assert np.array_equal(toff.cpu().numpy(), woff)
    assert np.array_equal(tnz.cpu().numpy(), wnz)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(8))
def test_transpose_and_permute_random_shapes(dev, orc, seed):
    """Seeded random shapes: rectangular, empty rows and columns, repeated
    columns, a few long rows; transpose and permute against the oracle."""
    rng = np.random.default_rng(1000 + seed)
    numrows = int(rng.integers(1, 6000))
    numcols = int(rng.choice([1, 2, 255, 256, 257, int(rng.integers(1, 1 << 17)), 1 << 20]))
    lens = rng.poisson(float(rng.choice([0.3, 3.0, 12.0])), numrows)
    lens[rng.integers(0, numrows, 3)] += int(rng.integers(0, 3000))  # a few long rows
    off = np.zeros(numrows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    nz = rng.integers(0, numcols, off[-1], dtype=np.int64)
    toff, tnz = B.spmat_transpose(numcols, _dev(off), _dev(nz))
    woff, wnz = orc.transpose_agi(numrows, numcols, off, nz)
    assert np.array_equal(toff.cpu().numpy(), woff) and np.array_equal(tnz.cpu().numpy(), wnz)
    # rectangular too: rperminv over the rows, cperminv over the columns
    rpi = rng.permutation(numrows).astype(np.int64)
    cpi = rng.permutation(numcols).astype(np.int64)
    poff, pnz = B.spmat_permute(numcols, _dev(off), _dev(nz), _dev(rpi), _dev(cpi))
    woff, wnz = orc.permute_agi(off, nz, rpi, cpi)
    assert np.array_equal(poff.cpu().numpy(), woff) and np.array_equal(pnz.cpu().numpy(), wnz)


@pytest.mark.gpu
@pytest.mark.parametrize("numcols", [3, 1 << 16, 1 << 20, 1 << 24])
def test_transpose_rows_across_tiles(dev, orc, numcols):
    """Enough nonzeros for many 8192-item tiles per chunk, with rows that run
    across tile and chunk boundaries: a stretch of 100000 empty rows, one row
    of 300000 nonzeros, single-nonzero rows, repeated columns.  numcols picks
    one pass (3), two (2^16: the last pass by column groups), three with a
    counted middle pass (2^20) and three with both later passes by groups
    (2^24)."""
    rng = np.random.default_rng(77 + numcols)
    numrows = 4_000_000
    lens = rng.poisson(5.0, numrows)
    lens[1_000_000:1_100_000] = 0
    lens[2_500_000] = 300_000
    lens[3_000_000:3_200_000] = 1
    off = np.zeros(numrows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    nz = rng.integers(0, numcols, off[-1], dtype=np.int64)
    toff, tnz = B.spmat_transpose(numcols, _dev(off), _dev(nz))
    woff, wnz = orc.transpose_agi(numrows, numcols, off, nz)
    assert np.array_equal(toff.cpu().numpy(), woff)
    assert np.array_equal(tnz.cpu().numpy(), wnz)


@pytest.mark.gpu
@pytest.mark.parametrize("numrows,mean", [(3_000_000, 5.0), (400_000, 5.0)])
def test_transpose_four_passes(dev, orc, numrows, mean):
    """2^25 + 3 columns: four radix passes (7, 7, 6, 6 bits) -- a counted
    middle pass (128 groups), a middle pass by groups (16384 groups, its
    prefix over columns modulo 2^20) and the last by groups.  The first case
    counts the columns in int32 inside the output array; the second has more
    than twice as many columns as nonzeros and counts them as int64."""
    numcols = (1 << 25) + 3
    rng = np.random.default_rng(numrows)
    lens = rng.poisson(mean, numrows)
    off = np.zeros(numrows + 1, np.int64)
    off[1:] = np.cumsum(lens)
    nz = rng.integers(0, numcols, off[-1], dtype=np.int64)
    toff, tnz = B.spmat_transpose(numcols, _dev(off), _dev(nz))
    woff, wnz = orc.transpose_agi(numrows, numcols, off, nz)
    assert np.array_equal(toff.cpu().numpy(), woff)
    assert np.array_equal(tnz.cpu().numpy(), wnz)


@pytest.mark.gpu
def test_transpose_in_place(dev, orc):
    """t_nonzero may be the matrix's own nonzero array (square matrices): the
    result equals the transpose into fresh arrays."""
    import torch
    n = 1 << 20
    off, nz = B.spmat_gen_er(n, 18 / n, 0, 3, seed=7)
    want_off, want_nz = B.spmat_transpose(n, off, nz)
    toff = torch.empty(n + 1, dtype=torch.int64, device=dev)
    B.check(B.lib().bale_spmat_transpose(n, n, nz.numel(), off.data_ptr(), nz.data_ptr(),
                                         toff.data_ptr(), nz.data_ptr(), None))
    assert torch.equal(toff, want_off) and torch.equal(nz, want_nz)


@pytest.mark.gpu
def test_transpose_er_and_back(dev, orc):
    n = 1 << 21
    off, nz = B.spmat_gen_er(n, 18 / n, 0, 3, seed=101894)
    toff, tnz = B.spmat_transpose(n, off, nz)
    woff, wnz = orc.transpose_agi(n, n, off.cpu().numpy(), nz.cpu().numpy())
    assert np.array_equal(toff.cpu().numpy(), woff) and np.array_equal(tnz.cpu().numpy(), wnz)
    # rows of an ER matrix are sorted, so transposing twice gives it back
    ooff, onz = B.spmat_transpose(n, toff, tnz)
    assert bool((ooff == off).all()) and bool((onz == nz).all())


@pytest.mark.gpu
def test_transpose_rejects_bad_columns(dev):
    off = _dev(np.array([0, 2, 3], np.int64))
    nz = _dev(np.array([0, 5, 1], np.int64))  # 5 is outside 3 columns
    with pytest.raises(B.BaleError) as e:
        B.spmat_transpose(3, off, nz)
    assert e.value.code == -2


@pytest.mark.gpu
# the last two take the column map through the library's gather (at least
# 2^24 nonzeros): by table region at 2^22 columns, one load per request at 2^21
@pytest.mark.parametrize("n,mean", [(1, 1), (777, 5), (100000, 9), (1 << 22, 5), (1 << 21, 12)])
def test_permute_matches_oracle(dev, orc, n, mean):
    rng = np.random.default_rng(n)
    lens = rng.poisson(mean, n)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(lens)
    nz = rng.integers(0, n, off[-1], dtype=np.int64)
    rpi, cpi = rng.permutation(n).astype(np.int64), rng.permutation(n).astype(np.int64)
    poff, pnz = B.spmat_permute(n, _dev(off), _dev(nz), _dev(rpi), _dev(cpi))
    woff, wnz = orc.permute_agi(off, nz, rpi, cpi)
    assert np.array_equal(poff.cpu().numpy(), woff)
    assert np.array_equal(pnz.cpu().numpy(), wnz)


@pytest.mark.gpu
def test_permute_reports_a_bad_column_in_a_large_matrix(dev):
    """The gather that maps the columns of a large matrix counts a column out
    of range like the small path does."""
    n = 1 << 22
    rng = np.random.default_rng(5)
    lens = rng.poisson(5, n)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(lens)
    nz = rng.integers(0, n, off[-1], dtype=np.int64)
    nz[12345] = n + 7
    rpi, cpi = rng.permutation(n).astype(np.int64), rng.permutation(n).astype(np.int64)
    with pytest.raises(B.BaleError) as e:
        B.spmat_permute(n, _dev(off), _dev(nz), _dev(rpi), _dev(cpi))
    assert e.value.code == -2


@pytest.mark.gpu
def test_permute_rejects_bad_permutation(dev):
    off = _dev(np.array([0, 1, 2], np.int64))
    nz = _dev(np.array([0, 1], np.int64))
    with pytest.raises(B.BaleError) as e:
        B.spmat_permute(2, off, nz, _dev(np.array([0, 0], np.int64)), _dev(np.array([1, 0], np.int64)))
    assert e.value.code == -2


@pytest.mark.gpu
def test_triangular_checks(dev, orc):
    off, nz = B.spmat_gen_er(3000, 0.01, 1, 2, seed=7)
    assert B.spmat_triangular(off, nz, True) == (0, 0)
    wrong, nopivot = B.spmat_triangular(off, nz, False)
    assert wrong == int(nz.numel()) - 3000 and nopivot == 0
    off, nz = B.spmat_gen_er(3000, 0.01, 0, 1, seed=7)
    assert B.spmat_triangular(off, nz, False) == (0, 3000)


@pytest.mark.gpu
def test_spmat_dropins(dev, orc):
    import ctypes as C
    L = B.lib()
    n, seed = 20000, 101894
    A = L.gen_erdos_renyi_graph_dist(n, 18 / n, 0, 3, seed)
    assert A, B.last_error()
    a = A.contents
    assert a.numrows == a.lnumrows == n and a.nnz == a.lnnz and a.offset == a.loffset
    off = np.empty(n + 1, np.int64)
    nz = np.empty(a.nnz, np.int64)
    L.bale_memget(off.ctypes.data, a.offset, off.nbytes)
    L.bale_memget(nz.ctypes.data, a.nonzero, nz.nbytes)
    woff, wnz = orc.gen_er(n, 18 / n, 0, 3, seed)
    assert np.array_equal(off, woff) and np.array_equal(nz, wnz)
    toff, tnz = orc.transpose_agi(n, n, off, nz)
    for name, extra in (("transpose_matrix_agi", ()), ("transpose_matrix_exstack", (1024,)),
                        ("transpose_matrix_exstack2", (1024,)), ("transpose_matrix_conveyor", ()),
                        ("transpose_matrix", ())):
        T = getattr(L, name)(A, *extra)
        assert T, B.last_error()
        t = T.contents
        got = np.empty(t.nnz, np.int64)
        L.bale_memget(got.ctypes.data, t.nonzero, got.nbytes)
        assert np.array_equal(got, tnz), name
        TT = L.transpose_matrix(T)
        assert L.compare_matrix(TT, A) == 0  # rows of A are sorted
        L.clear_matrix(TT)
        L.clear_matrix(T)
    # permutations from rand_permp, as the toposort driver makes them
    rpi_p, cpi_p = L.rand_permp(n, 1230), L.rand_permp(n, 45)
    rpi, cpi = np.empty(n, np.int64), np.empty(n, np.int64)
    L.bale_memget(rpi.ctypes.data, rpi_p, rpi.nbytes)
    L.bale_memget(cpi.ctypes.data, cpi_p, cpi.nbytes)
    poff, pnz = orc.permute_agi(off, nz, rpi, cpi)
    for name, extra in (("permute_matrix_agi", ()), ("permute_matrix_exstack", (1024,)),
                        ("permute_matrix_exstack2", (1024,)), ("permute_matrix_conveyor", ()),
                        ("permute_matrix", ())):
        P = getattr(L, name)(A, rpi_p, cpi_p, *extra)
        assert P, B.last_error()
        pp = P.contents
        got_off = np.empty(n + 1, np.int64)
        got = np.empty(pp.nnz, np.int64)
        L.bale_memget(got_off.ctypes.data, pp.offset, got_off.nbytes)
        L.bale_memget(got.ctypes.data, pp.nonzero, got.nbytes)
        assert np.array_equal(got_off, poff) and np.array_equal(got, pnz), name
        L.clear_matrix(P)
    # host arrays are accepted too
    h = B.SparseMat(0, 3, 3, 3, 3, 3, off.ctypes.data, off.ctypes.data, nz.ctypes.data, nz.ctypes.data)
    hoff = np.array([0, 1, 3, 3], np.int64)
    hnz = np.array([2, 0, 1], np.int64)
    h.offset = h.loffset = hoff.ctypes.data
    h.nonzero = h.lnonzero = hnz.ctypes.data
    T = L.transpose_matrix_agi(C.byref(h))
    assert T, B.last_error()
    got_off, got = np.empty(4, np.int64), np.empty(3, np.int64)
    L.bale_memget(got_off.ctypes.data, T.contents.offset, 32)
    L.bale_memget(got.ctypes.data, T.contents.nonzero, 24)
    assert list(got_off) == [0, 1, 2, 3] and list(got) == [1, 1, 0]
    L.clear_matrix(T)
    U = L.gen_erdos_renyi_graph_dist(2000, 0.01, 1, 2, 3)
    assert L.is_upper_triangular(U, 1) == 1 and L.is_lower_triangular(U, 1) == 0
    L.clear_matrix(U)
    L.clear_matrix(A)
    L.bale_all_free(rpi_p)
    L.bale_all_free(cpi_p)
    assert not L.gen_erdos_renyi_graph_dist(10, 0.1, 0, 0, 1)  # mode 0 is not provided
