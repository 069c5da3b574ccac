"""ctypes bindings to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product (bale_old_amd, libbale_hip.so) never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None


class Instance(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dist", C.c_int32), ("table_len", C.c_int64),
                ("nops", C.c_int64), ("s", C.c_double), ("seed", C.c_uint64)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, I, U = C.c_void_p, C.c_int64, C.c_uint64
        sig = {
            "or_splitmix64_next": (U, [C.POINTER(C.c_uint64)]),
            "or_splitmix64_at": (U, [U, U]),
            "or_splitmix64": (U, [U]),
            "or_reduce": (U, [U, U]),
            "or_perm": (U, [U, U, U]),
            "or_gen_uniform": (None, [P, I, I, U]),
            "or_gen_index": (None, [P, I, I, C.c_int, C.c_double, I, U]),
            "or_zipf_pmf": (C.c_double, [I, I, C.c_double]),
            "or_fill_splitmix": (None, [P, I]),
            "or_fill_neg": (None, [P, I]),
            "or_pack": (None, [P, P, I, I]),
            "or_gen_reference_rand": (None, [P, I, I, C.c_uint]),
            "or_histo_agi": (None, [P, I, P]),
            "or_histo_agi_i32": (None, [P, I, P]),
            "or_histo_owner": (C.c_int, [P, I, P, I, I]),
            "or_histo_agi_mt": (C.c_double, [P, I, P, C.c_int]),
            "or_histo_owner_mt": (C.c_double, [P, I, P, C.c_int]),
            "or_ig_owner_mt": (C.c_double, [P, P, I, P, C.c_int]),
            "or_ig_agi": (None, [P, P, I, P]),
            "or_ig_exstack": (C.c_int, [P, P, I, P, I, I]),
            "or_ig_conveyor": (C.c_int, [P, P, I, P, I, I]),
            "or_ig_agi_mt": (C.c_double, [P, P, I, P, C.c_int]),
            "or_batch_params": (None, [P, I, U, C.c_int, C.c_int, C.c_int, C.c_int]),
            "or_digest": (U, [P, I]),
            "or_convey_deliver": (I, [P, P, I, I, I, P, P]),
            "or_srand48_state": (U, [I]),
            "or_lrand48_seq": (None, [U, I, P]),
            "or_glibc_lrand48": (None, [I, I, P]),
            "or_randperm_agi": (I, [P, I, I]),
            "or_randperm_agi_opt": (None, [P, I, I]),
            "or_is_perm": (C.c_int, [P, I]),
            "or_det_log": (C.c_double, [C.c_double]),
            "or_gen_er": (I, [I, C.c_double, C.c_int, C.c_int, U, P, P]),
            "or_transpose_agi": (None, [I, I, P, P, P, P]),
            "or_permute_agi": (None, [I, P, P, P, P, P, P]),
            "or_toposort_agi": (I, [I, P, P, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def splitmix64_at(seed: int, i: int) -> int:
    return lib().or_splitmix64_at(seed, i)


def splitmix64(j: int) -> int:
    return lib().or_splitmix64(j)


def reduce(x: int, n: int) -> int:
    return lib().or_reduce(x, n)


def perm(x: int, n: int, key: int) -> int:
    return lib().or_perm(x, n, key)


def gen_uniform(n: int, rng: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().or_gen_uniform(_p(out), n, rng, seed)
    return out


def gen_index(n: int, nbins: int, dist: int = 0, s: float = 1.0, run_mean: int = 0, seed: int = 1) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().or_gen_index(_p(out), n, nbins, dist, s, run_mean, seed)
    return out


def zipf_pmf(k: int, n: int, s: float) -> float:
    return lib().or_zipf_pmf(k, n, s)


def fill_splitmix(n: int) -> np.ndarray:
    t = np.empty(n, np.int64)
    lib().or_fill_splitmix(_p(t), n)
    return t


def fill_neg(n: int) -> np.ndarray:
    t = np.empty(n, np.int64)
    lib().or_fill_neg(_p(t), n)
    return t


def pack(idx: np.ndarray, P: int = 1) -> np.ndarray:
    out = np.empty_like(idx)
    lib().or_pack(_p(out), _p(idx), idx.size, P)
    return out


def reference_rand(n: int, rng: int, seed0: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().or_gen_reference_rand(_p(out), n, rng, seed0)
    return out


def histo_agi(index: np.ndarray, counts: np.ndarray) -> np.ndarray:
    if counts.dtype == np.int64:
        lib().or_histo_agi(_p(index), index.size, _p(counts))
    else:
        lib().or_histo_agi_i32(_p(index), index.size, _p(counts))
    return counts


def histo_owner(pck: np.ndarray, lcounts: np.ndarray, P: int) -> int:
    lnum = lcounts.size // P
    return lib().or_histo_owner(_p(pck), pck.size, _p(lcounts), lnum, P)


def histo_agi_mt(index: np.ndarray, counts: np.ndarray, nthreads: int) -> float:
    return lib().or_histo_agi_mt(_p(index), index.size, _p(counts), nthreads)


def ig_agi(index: np.ndarray, table: np.ndarray) -> np.ndarray:
    tgt = np.empty(index.size, np.int64)
    lib().or_ig_agi(_p(tgt), _p(index), index.size, _p(table))
    return tgt


def ig_exstack(pck: np.ndarray, ltables: np.ndarray, P: int) -> np.ndarray:
    tgt = np.zeros(pck.size, np.int64)
    rc = lib().or_ig_exstack(_p(tgt), _p(pck), pck.size, _p(ltables), ltables.size // P, P)
    assert rc == 0
    return tgt


def ig_conveyor(pck: np.ndarray, ltables: np.ndarray, P: int) -> np.ndarray:
    tgt = np.zeros(pck.size, np.int64)
    rc = lib().or_ig_conveyor(_p(tgt), _p(pck), pck.size, _p(ltables), ltables.size // P, P)
    assert rc == 0
    return tgt


def histo_owner_mt(index: np.ndarray, counts: np.ndarray, nthreads: int) -> float:
    """Aggregate-by-owner histogram over nthreads PEs; average seconds per PE."""
    return lib().or_histo_owner_mt(_p(index), index.size, _p(counts), nthreads)


def ig_owner_mt(index: np.ndarray, table: np.ndarray, tgt: np.ndarray, nthreads: int) -> float:
    """Request/reply gather by owner over nthreads PEs; average seconds per PE."""
    return lib().or_ig_owner_mt(_p(tgt), _p(index), index.size, _p(table), nthreads)


def ig_agi_mt(index: np.ndarray, table: np.ndarray, tgt: np.ndarray, nthreads: int) -> float:
    return lib().or_ig_agi_mt(_p(tgt), _p(index), index.size, _p(table), nthreads)


def batch_params(count: int, seed: int, lo_tab=10, hi_tab=24, lo_ops=12, hi_ops=22) -> list[dict]:
    arr = (Instance * count)()
    lib().or_batch_params(C.addressof(arr), count, seed, lo_tab, hi_tab, lo_ops, hi_ops)
    return [{"kind": x.kind, "dist": x.dist, "table_len": x.table_len, "nops": x.nops, "s": x.s,
             "seed": x.seed} for x in arr]


def convey_deliver(items: np.ndarray, pes: np.ndarray, P: int) -> tuple[np.ndarray, np.ndarray, int]:
    """(delivered items, starts[P+1], dropped) for items (n rows of bytes) sent to pes."""
    items = np.ascontiguousarray(items)
    pes = np.ascontiguousarray(pes, dtype=np.int64)
    n = pes.size
    ib = items.nbytes // n if n else 0
    out = np.zeros_like(items)
    starts = np.zeros(P + 1, dtype=np.int64)
    dropped = lib().or_convey_deliver(_p(items), _p(pes), n, ib, P, _p(out), _p(starts))
    return out[: starts[P]], starts, dropped


def digest(v: np.ndarray) -> int:
    return lib().or_digest(_p(np.ascontiguousarray(v, dtype=np.int64)), v.size)


def srand48_state(seed: int) -> int:
    return lib().or_srand48_state(seed)


def lrand48_seq(state: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().or_lrand48_seq(state, n, _p(out))
    return out


def glibc_lrand48(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().or_glibc_lrand48(seed, n, _p(out))
    return out


def randperm_agi(N: int, seed: int) -> tuple[np.ndarray, int]:
    """(permutation, darts thrown) of rand_permp_agi at one PE."""
    perm = np.empty(max(N, 1), np.int64)
    darts = lib().or_randperm_agi(_p(perm), N, seed)
    assert darts >= 0
    return perm[:N], darts


def randperm_agi_opt(N: int, seed: int) -> np.ndarray:
    """rand_permp_agi_opt at one PE: Fisher-Yates after N generator draws."""
    perm = np.empty(max(N, 1), np.int64)
    lib().or_randperm_agi_opt(_p(perm), N, seed)
    return perm[:N]


def is_perm(perm: np.ndarray, N: int) -> bool:
    perm = np.ascontiguousarray(perm, dtype=np.int64)
    return bool(lib().or_is_perm(_p(perm), N)) if perm.size == N else False


def det_log(x: float) -> float:
    return lib().or_det_log(x)


def gen_er(n: int, p: float, unit_diag: int = 0, mode: int = 3, seed: int = 1):
    """(offset, nonzero) of the Erdos-Renyi matrix the library generates."""
    off = np.empty(n + 1, np.int64)
    nnz = lib().or_gen_er(n, p, unit_diag, mode, seed, _p(off), None)
    nz = np.empty(max(nnz, 1), np.int64)
    lib().or_gen_er(n, p, unit_diag, mode, seed, _p(off), _p(nz))
    return off, nz[:nnz]


def transpose_agi(numrows: int, numcols: int, off: np.ndarray, nz: np.ndarray):
    off = np.ascontiguousarray(off, dtype=np.int64)
    nz = np.ascontiguousarray(nz, dtype=np.int64)
    toff = np.empty(numcols + 1, np.int64)
    tnz = np.empty(max(nz.size, 1), np.int64)
    lib().or_transpose_agi(numrows, numcols, _p(off), _p(nz), _p(toff), _p(tnz))
    return toff, tnz[: nz.size]


def permute_agi(off: np.ndarray, nz: np.ndarray, rperminv: np.ndarray, cperminv: np.ndarray):
    off = np.ascontiguousarray(off, dtype=np.int64)
    nz = np.ascontiguousarray(nz, dtype=np.int64)
    rperminv = np.ascontiguousarray(rperminv, dtype=np.int64)
    cperminv = np.ascontiguousarray(cperminv, dtype=np.int64)
    numrows = off.size - 1
    poff = np.empty(numrows + 1, np.int64)
    pnz = np.empty(max(nz.size, 1), np.int64)
    lib().or_permute_agi(numrows, _p(off), _p(nz), _p(rperminv), _p(cperminv), _p(poff), _p(pnz))
    return poff, pnz[: nz.size]


def toposort_agi(off: np.ndarray, nz: np.ndarray, toff: np.ndarray, tnz: np.ndarray):
    """(rperm, cperm, pivots, levels) of toposort_matrix_agi at one PE."""
    n = off.size - 1
    rperm = np.full(max(n, 1), -1, np.int64)
    cperm = np.full(max(n, 1), -1, np.int64)
    lev = np.zeros(1, np.int64)
    piv = lib().or_toposort_agi(n, _p(np.ascontiguousarray(off, dtype=np.int64)),
                                _p(np.ascontiguousarray(nz, dtype=np.int64)),
                                _p(np.ascontiguousarray(toff, dtype=np.int64)),
                                _p(np.ascontiguousarray(tnz, dtype=np.int64)), _p(rperm), _p(cperm),
                                _p(lev))
    return rperm[:n], cperm[:n], piv, int(lev[0])

