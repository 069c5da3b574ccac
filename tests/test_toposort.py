"""Toposort of a morally upper-triangular matrix (toposort_matrix_agi,
apps/topo_src/toposort_agi.upc:53-166).

CPU: the oracle pinned -- its permutations make the matrix upper triangular
with a full diagonal (the reference driver's own check), and equal those of an
independent Python restatement of the queue (a deque, levels kept apart);
level counts of a chain and of a diagonal.  GPU: the library's permutations
and level count equal to the oracle's, through the native API and the
drop-ins, on the reference driver's input pipeline.
"""
from collections import deque

import numpy as np
import pytest

import bale_old_amd as B


def make_csr(rows):
    off = np.zeros(len(rows) + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    nz = np.array([c for r in rows for c in r], np.int64)
    return off, nz


def toposort_input(orc, n, nz_per_row, seed, rseed=1230, cseed=45):
    """The driver's pipeline (apps/topo_src/toposort.upc:176-213): an upper
    triangular ER matrix with a full diagonal, rows and columns permuted."""
    p = min(1.0, 2.0 * nz_per_row / (n - 1)) if n > 1 else 0.0
    off, nz = orc.gen_er(n, p, 1, 2, seed)
    rpi, _ = orc.randperm_agi(n, rseed)
    cpi, _ = orc.randperm_agi(n, cseed)
    moff, mnz = orc.permute_agi(off, nz, rpi, cpi)
    toff, tnz = orc.transpose_agi(n, n, moff, mnz)
    return moff, mnz, toff, tnz


def py_toposort(off, nz, toff, tnz):
    """The reference's loop with Python lists (one PE)."""
    n = off.size - 1
    cnt = [int(off[i + 1] - off[i]) for i in range(n)]
    rsum = [int(nz[off[i]:off[i + 1]].sum()) for i in range(n)]
    level = deque(i for i in range(n) if cnt[i] == 1)
    rperm, cperm = [-1] * n, [-1] * n
    pos = levels = 0
    while level:
        levels += 1
        nxt = deque()
        while level:
            r = level.popleft()
            c = rsum[r]
            rperm[r] = cperm[c] = n - 1 - pos
            pos += 1
            for s in tnz[toff[c]:toff[c + 1]]:
                cnt[s] -= 1
                rsum[s] -= c
                if cnt[s] == 1:
                    nxt.append(int(s))
        level = nxt
    return np.array(rperm, np.int64), np.array(cperm, np.int64), pos, levels


def upper_with_diagonal(orc, off, nz, rperm, cperm):
    poff, pnz = orc.permute_agi(off, nz, rperm, cperm)
    for r in range(off.size - 1):
        row = pnz[poff[r]:poff[r + 1]]
        if not (np.all(row >= r) and r in row):
            return False
    return True


# ---- CPU: the oracle pinned ------------------------------------------------------------
@pytest.mark.parametrize("n,z,seed", [(1, 1, 1), (2, 1, 3), (50, 3, 5), (2000, 10, 7), (5000, 35, 5)])
def test_oracle_matches_python_and_triangularizes(orc, n, z, seed):
    off, nz, toff, tnz = toposort_input(orc, n, z, seed)
    rperm, cperm, piv, lev = orc.toposort_agi(off, nz, toff, tnz)
    prp, pcp, ppiv, plev = py_toposort(off, nz, toff, tnz)
    assert piv == ppiv == n and lev == plev
    assert np.array_equal(rperm, prp) and np.array_equal(cperm, pcp)
    assert orc.is_perm(rperm, n) and orc.is_perm(cperm, n)
    assert upper_with_diagonal(orc, off, nz, rperm, cperm)


def test_oracle_levels_of_chain_and_diagonal(orc):
    n = 300
    # a chain: row i holds columns i and i+1, so one row frees the next
    off, nz = make_csr([[i, i + 1] if i + 1 < n else [i] for i in range(n)])
    toff, tnz = orc.transpose_agi(n, n, off, nz)
    rperm, cperm, piv, lev = orc.toposort_agi(off, nz, toff, tnz)
    assert piv == n and lev == n
    assert list(rperm) == list(range(n)) and list(cperm) == list(range(n))
    # the diagonal: one level, rows in order take n-1, n-2, ...
    off, nz = make_csr([[i] for i in range(n)])
    rperm, cperm, piv, lev = orc.toposort_agi(off, nz, off, nz)
    assert piv == n and lev == 1 and list(rperm) == list(range(n - 1, -1, -1))


def test_oracle_stops_on_a_non_triangular_matrix(orc):
    off, nz = make_csr([[0, 1], [0, 1]])
    rperm, cperm, piv, lev = orc.toposort_agi(off, nz, off, nz)
    assert piv == 0 and lev == 0


def test_oracle_pivot_rows_empty_themselves(orc):
    # a pivot row is in its own column, so its count goes to 0; the other row
    # of that column is freed (3 x 3, one row per level)
    off, nz = make_csr([[1], [0, 1], [2, 0]])
    toff, tnz = orc.transpose_agi(3, 3, off, nz)
    rperm, cperm, piv, lev = orc.toposort_agi(off, nz, toff, tnz)
    assert piv == 3 and lev == 3
    assert list(rperm) == [2, 1, 0] and list(cperm) == [1, 2, 0]


# ---- GPU -----------------------------------------------------------------------------------
def _dev(t):
    import torch
    return torch.from_numpy(np.ascontiguousarray(t)).to("cuda:0")


@pytest.mark.gpu
@pytest.mark.parametrize("n,z,seed", [(1, 1, 1), (2, 1, 3), (1000, 3, 9), (100000, 35, 5),
                                      (1 << 20, 35, 5)])
def test_toposort_matches_oracle(dev, orc, n, z, seed):
    off, nz, toff, tnz = toposort_input(orc, n, z, seed)
    want_r, want_c, piv, lev = orc.toposort_agi(off, nz, toff, tnz)
    assert piv == n
    rperm, cperm, levels = B.toposort(_dev(off), _dev(nz), _dev(toff), _dev(tnz))
    assert levels == lev
    assert np.array_equal(rperm.cpu().numpy(), want_r)
    assert np.array_equal(cperm.cpu().numpy(), want_c)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_toposort_random_inputs(dev, orc, seed):
    """Seeded random sizes and densities, from a handful of rows to 2^16, as
    sparse as a chain and as dense as 200 per row."""
    rng = np.random.default_rng(500 + seed)
    n = int(rng.integers(2, 1 << 16))
    z = float(rng.choice([0.5, 2.0, 10.0, 35.0, 200.0]))
    off, nz, toff, tnz = toposort_input(orc, n, z, int(rng.integers(1, 1 << 30)),
                                        int(rng.integers(1, 1 << 20)), int(rng.integers(1, 1 << 20)))
    want_r, want_c, piv, lev = orc.toposort_agi(off, nz, toff, tnz)
    assert piv == n
    rperm, cperm, levels = B.toposort(_dev(off), _dev(nz), _dev(toff), _dev(tnz))
    assert levels == lev
    assert np.array_equal(rperm.cpu().numpy(), want_r) and np.array_equal(cperm.cpu().numpy(), want_c)


@pytest.mark.gpu
def test_toposort_separate_counters(dev, orc, monkeypatch):
    """The path for matrices whose row counts and column sums do not share a
    64-bit word, forced by BALE_TOPOSORT_SEPARATE."""
    monkeypatch.setenv("BALE_TOPOSORT_SEPARATE", "1")
    off, nz, toff, tnz = toposort_input(orc, 100000, 35, 5)
    want_r, want_c, piv, lev = orc.toposort_agi(off, nz, toff, tnz)
    rperm, cperm, levels = B.toposort(_dev(off), _dev(nz), _dev(toff), _dev(tnz))
    assert levels == lev
    assert np.array_equal(rperm.cpu().numpy(), want_r) and np.array_equal(cperm.cpu().numpy(), want_c)


@pytest.mark.gpu
def test_toposort_chain_has_one_level_per_row(dev, orc):
    n = 1000  # 1000 levels: many batches of launches
    rows = [[i, i + 1] if i + 1 < n else [i] for i in range(n)]
    rng = np.random.default_rng(2)
    rpi, cpi = rng.permutation(n).astype(np.int64), rng.permutation(n).astype(np.int64)
    off, nz = make_csr(rows)
    moff, mnz = orc.permute_agi(off, nz, rpi, cpi)
    toff, tnz = orc.transpose_agi(n, n, moff, mnz)
    want_r, want_c, piv, lev = orc.toposort_agi(moff, mnz, toff, tnz)
    rperm, cperm, levels = B.toposort(_dev(moff), _dev(mnz), _dev(toff), _dev(tnz))
    assert levels == lev == n
    assert np.array_equal(rperm.cpu().numpy(), want_r) and np.array_equal(cperm.cpu().numpy(), want_c)


@pytest.mark.gpu
def test_toposort_small_by_hand(dev, orc):
    off, nz = make_csr([[1], [0, 1], [2, 0]])
    toff, tnz = orc.transpose_agi(3, 3, off, nz)
    rperm, cperm, levels = B.toposort(_dev(off), _dev(nz), _dev(toff), _dev(tnz))
    assert levels == 3
    assert rperm.tolist() == [2, 1, 0] and cperm.tolist() == [1, 2, 0]


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [[[0, 1], [0, 1]], [[0], [0]], [[0], [0], [1, 2]]])
def test_toposort_reports_a_non_triangular_matrix(dev, orc, rows):
    off, nz = make_csr(rows)
    n = off.size - 1
    toff, tnz = orc.transpose_agi(n, n, off, nz)
    with pytest.raises(B.BaleError) as e:
        B.toposort(_dev(off), _dev(nz), _dev(toff), _dev(tnz))
    assert e.value.code == -1


@pytest.mark.gpu
def test_toposort_dropins_on_the_driver_pipeline(dev, orc):
    """gen_erdos_renyi_graph_dist -> rand_permp -> permute_matrix ->
    transpose_matrix -> toposort_matrix_* -> check_is_triangle, all through
    the drop-ins (apps/topo_src/toposort.upc:135-213, 283-335)."""
    import ctypes as C
    L = B.lib()
    n = 50000
    p = 2.0 * 35 / (n - 1)
    omat = L.gen_erdos_renyi_graph_dist(n, p, 1, 2, 5)
    assert omat and L.is_upper_triangular(omat, 1) == 1
    rpi, cpi = L.rand_permp(n, 1230), L.rand_permp(n, 45)
    mat = L.permute_matrix(omat, rpi, cpi)
    tmat = L.transpose_matrix(mat)
    assert mat and tmat, B.last_error()
    m, t = mat.contents, tmat.contents
    get = lambda ptr, k: (lambda a: (L.bale_memget(a.ctypes.data, ptr, a.nbytes), a)[1])(np.empty(k, np.int64))
    moff, mnz = get(m.offset, n + 1), get(m.nonzero, m.nnz)
    toff, tnz = get(t.offset, n + 1), get(t.nonzero, t.nnz)
    want_r, want_c, piv, _ = orc.toposort_agi(moff, mnz, toff, tnz)
    assert piv == n
    rperm = L.bale_all_alloc(n, 8)
    cperm = L.bale_all_alloc(n, 8)
    for name in ("toposort_matrix_agi", "toposort_matrix_exstack", "toposort_matrix_exstack2",
                 "toposort_matrix_convey", "toposort_matrix_oo", "toposort_matrix_cooler",
                 "toposort_matrix_exstack_orig"):
        secs = getattr(L, name)(rperm, cperm, mat, tmat)
        assert secs > 0, (name, B.last_error())
        assert np.array_equal(get(rperm, n), want_r) and np.array_equal(get(cperm, n), want_c), name
        # the driver's check_is_triangle (toposort.upc:147-167)
        assert L.is_perm(rperm, n) == 1 and L.is_perm(cperm, n) == 1
        mat2 = L.permute_matrix(mat, rperm, cperm)
        assert L.is_upper_triangular(mat2, 1) == 1
        L.clear_matrix(mat2)
    # host arrays for the permutations work too (staged and copied back)
    hr, hc = np.full(n, -7, np.int64), np.full(n, -7, np.int64)
    assert L.toposort_matrix_agi(hr.ctypes.data, hc.ctypes.data, mat, tmat) > 0, B.last_error()
    assert np.array_equal(hr, want_r) and np.array_equal(hc, want_c)
    assert L.is_perm(hr.ctypes.data, n) == 1
    for x in (omat, mat, tmat):
        L.clear_matrix(x)
    for ptr in (rpi, cpi, rperm, cperm):
        L.bale_all_free(ptr)
