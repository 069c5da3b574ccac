"""The packed 16-bit LDS tile's accounting (bale_old_amd/csrc/pk_count.h),
replayed on the host.

The bucketed histogram's apply pass keeps two 16-bit counters per 32-bit LDS
word for 2^16-bin buckets.  Its counts must equal the reference's
(apps/histo_src/histo_agi.upc:61-66, checked by histo.upc:208-227) whatever
order the lanes' atomics land in and however late a lane accounts for the
value its atomic returned.  tests/pk_replay.cpp runs schedules through the
product's own pk_owed / pk_read; the expected counts are plain sums.

The earlier rule (move 1024 off a counter each time an add carries it across
a multiple of 1024, as a second atomic issued after the first returned) is
restated here in Python to show what such schedules do to it: when enough
corrections are still owed the half wraps and the count is lost.
"""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    so = tmp_path_factory.mktemp("pk") / "libpkreplay.so"
    subprocess.run([cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-D__host__=", "-D__device__=",
                    f"-I{ROOT / 'bale_old_amd' / 'csrc'}", str(ROOT / "tests" / "pk_replay.cpp"),
                    "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    P = C.c_void_p
    lib.pk_replay.argtypes = [P, P, P, C.c_int64, C.c_uint32, P, P]
    lib.pk_replay.restype = C.c_int

    def run(bins, incs, account_at, width, init=None):
        bins = np.ascontiguousarray(bins, np.uint32)
        incs = np.ascontiguousarray(incs, np.uint32)
        acc = np.ascontiguousarray(account_at, np.int64)
        words = np.zeros((width + 1) // 2, np.uint32) if init is None else np.ascontiguousarray(init, np.uint32)
        out = np.zeros(width, np.uint64)
        rc = lib.pk_replay(bins.ctypes.data, incs.ctypes.data, acc.ctypes.data, len(bins), width,
                           words.ctypes.data, out.ctypes.data)
        assert rc == 0, rc
        return out
    return run


def expected(bins, incs, width, init=None):
    want = np.bincount(np.asarray(bins, np.int64), weights=np.asarray(incs, np.float64),
                       minlength=width).astype(np.uint64)
    if init is not None:
        init = np.asarray(init, np.uint64)
        want[0::2] += init & 0xffff
        want[1::2] += (init >> 16)[: width // 2]
    return want


def old_rule(bins, incs, account_at, width, spill=1024):
    """The earlier kSpill rule: a lane whose add carried its counter across a
    multiple of `spill` later subtracts `spill` from the word (a second atomic)
    and records a spill; count = half + spill * spills."""
    word = [0] * ((width + 1) // 2)
    spills = [0] * width
    pending = {}
    for k, (b, i, a) in enumerate(zip(bins, incs, account_at)):
        w = b >> 1
        sft = 16 * (b & 1)
        old = word[w]
        word[w] = (old + (int(i) << sft)) & 0xffffffff
        pending.setdefault(int(a), []).append((int(b), int(i), old))
        for (pb, pi, pold) in pending.pop(k, []):
            ps = 16 * (pb & 1)
            if (((pold >> ps) & 0xffff) & (spill - 1)) + pi >= spill:
                word[pb >> 1] = (word[pb >> 1] - (spill << ps)) & 0xffffffff
                spills[pb] += 1
    return np.array([((word[b >> 1] >> (16 * (b & 1))) & 0xffff) + spill * spills[b]
                     for b in range(width)], np.uint64)


def test_in_order_accounting_is_exact(replay):
    rng = np.random.default_rng(1)
    n = 300_000
    bins = rng.integers(0, 2, n)
    incs = rng.integers(1, 9, n)
    got = replay(bins, incs, np.arange(n), 2)
    assert np.array_equal(got, expected(bins, incs, 2))


@pytest.mark.parametrize("seed", range(4))
def test_random_late_accounting_is_exact(replay, seed):
    # lanes account anywhere from at once to thousands of adds later
    rng = np.random.default_rng(100 + seed)
    n, width = 400_000, 6
    bins = rng.choice(width, n, p=[0.45, 0.45, 0.04, 0.03, 0.02, 0.01])
    incs = rng.integers(1, 9, n)
    acc = np.minimum(np.arange(n) + rng.geometric(1e-3, n) - 1, n - 1)
    got = replay(bins, incs, acc, width)
    assert np.array_equal(got, expected(bins, incs, width))


def test_every_correction_deferred_to_the_end(replay):
    # 2^17 + 5 adds on the low bin and 2^16 + 3 on the high bin, no lane
    # accounts until every add landed
    bins = np.r_[np.zeros(1 << 17, int), np.ones(1 << 16, int), [0] * 5, [1] * 3]
    incs = np.ones(len(bins), int)
    acc = np.full(len(bins), len(bins) - 1)
    assert np.array_equal(replay(bins, incs, acc, 2), expected(bins, incs, 2))


def test_high_half_crossing_while_a_carry_is_owed(replay):
    # The schedule that breaks the earlier rule: the high bin sits at 1022;
    # 2^16 low adds land with their corrections owed, so the low half wraps
    # and carries 1 into the high half; a high add then sees 1023 and spills;
    # the low corrections land (borrowing the carry back); the next high add
    # sees 1023 again and spills a second time: two spills for 1024 adds.
    n_hi0, n_lo = 1022, 1 << 16
    bins = np.r_[np.ones(n_hi0, int), np.zeros(n_lo, int), [1, 1]]
    incs = np.ones(len(bins), int)
    p1 = n_hi0 + n_lo  # the first of the last two high adds
    acc = np.arange(len(bins))
    acc[n_hi0:p1] = p1
    want = expected(bins, incs, 2)
    assert want.tolist() == [n_lo, n_hi0 + 2]
    assert np.array_equal(replay(bins, incs, acc, 2), want)
    got_old = old_rule(bins, incs, acc, 2)
    assert got_old[0] == want[0] and got_old[1] == want[1] + (1 << 16)


def test_carry_into_a_full_high_half(replay):
    # high half at 0xffff when the low half wraps: the carry wraps it too
    init = np.array([0xffff_fff0], np.uint32)  # high 0xffff, low 0xfff0
    bins = np.array([0] * 40 + [1] * 3 + [0] * 70000)
    incs = np.ones(len(bins), int)
    for acc in (np.arange(len(bins)), np.full(len(bins), len(bins) - 1)):
        assert np.array_equal(replay(bins, incs, acc, 2, init), expected(bins, incs, 2, init))


def test_runs_merged_into_larger_adds(replay):
    # the RUNS path adds up to 8 at once; wraps land mid-add
    rng = np.random.default_rng(7)
    n = 200_000
    bins = rng.integers(0, 4, n)
    incs = rng.integers(1, 9, n)
    acc = np.minimum(np.arange(n) + rng.integers(0, 50_000, n), n - 1)
    assert np.array_equal(replay(bins, incs, acc, 4), expected(bins, incs, 4))


def test_odd_width_last_high_half_has_no_bin(replay):
    # width 3: word 1's high half is no bin; carries into it are not counted
    bins = np.array([2] * 200_000 + [0] * 70_000 + [1] * 5)
    incs = np.ones(len(bins), int)
    acc = np.full(len(bins), len(bins) - 1)
    assert np.array_equal(replay(bins, incs, acc, 3), expected(bins, incs, 3))


def test_old_rule_with_prompt_corrections_was_exact():
    # the earlier rule held when every correction landed before the next add
    rng = np.random.default_rng(3)
    n = 100_000
    bins = rng.integers(0, 2, n)
    incs = rng.integers(1, 9, n)
    assert np.array_equal(old_rule(bins, incs, np.arange(n), 2), expected(bins, incs, 2))
