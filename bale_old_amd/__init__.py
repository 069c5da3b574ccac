"""bale_old_amd -- MI355X (gfx950) histogram scatter-add and index gather.

The product is ``libbale_hip.so`` (C-ABI declared in ``include/bale_hip.h``).
This module only loads it with ctypes and passes torch tensors' device
pointers and streams through; it never computes a result itself.  If the
library is missing, :func:`lib` raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libbale_hip.so"

# error codes (include/bale_hip.h)
OK, ERR_ARG, ERR_RANGE, ERR_HIP, ERR_ALLOC, ERR_PE, ERR_NODEV = 0, -1, -2, -3, -4, -5, -6
# strategies / dtypes / formats / merge modes
AUTO, DIRECT, LDS, BUCKET = 0, 1, 2, 3
I64, I32 = 0, 1
PLAIN, PACKED = 0, 1
MERGE_NONE, MERGE_RUNS, MERGE_MATCH = 0, 1, 2

STRATEGY_NAMES = {AUTO: "auto", DIRECT: "direct", LDS: "lds", BUCKET: "bucket"}


class BaleConfig(C.Structure):
    _fields_ = [
        ("strategy", C.c_int32),
        ("block", C.c_int32),
        ("unroll", C.c_int32),
        ("tile_log2", C.c_int32),
        ("merge", C.c_int32),
        ("groups_per_cu", C.c_int32),
        ("hot", C.c_int32),
        ("task_log2", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}

    def __repr__(self) -> str:
        return f"BaleConfig({self.as_dict()})"


def config(strategy=AUTO, block=0, unroll=0, tile_log2=0, merge=-1, groups_per_cu=0, hot=0,
           task_log2=0) -> BaleConfig:
    return BaleConfig(strategy, block, unroll, tile_log2, merge, groups_per_cu, hot, task_log2)


class Instance(C.Structure):
    """bale_instance: one histogram (kind 0) or gather (kind 1) of a batch."""
    _fields_ = [
        ("kind", C.c_int32),
        ("dtype", C.c_int32),
        ("n", C.c_int64),
        ("table_len", C.c_int64),
        ("index", C.c_void_p),
        ("table", C.c_void_p),
        ("tgt", C.c_void_p),
        ("cfg", BaleConfig),
    ]


class SparseMat(C.Structure):
    """The reference's sparsemat_t (spmat/spmat.h:73-86), field for field."""
    _fields_ = [(n, C.c_int64) for n in ("local", "numrows", "lnumrows", "numcols", "nnz", "lnnz")] + [
        (n, C.c_void_p) for n in ("offset", "loffset", "nonzero", "lnonzero")]


class BaleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bale error {code}: {msg}")
        self.code = code


_lib = None

P64 = C.POINTER(C.c_int64)
VP = C.c_void_p
I64T = C.c_int64
DROPIN_HISTO = ["histo_agi"]
DROPIN_HISTO_BUF = ["histo_exstack", "histo_exstack2", "histo_exstack2_goto", "histo_exstack2_cyclic",
                    "histo_exstack_function", "histo_exstack2_function"]
DROPIN_IG_BUF = ["ig_exstack", "ig_exstack2", "ig_exstack2_cyclic", "ig_exstack2_goto",
                 "ig_exstack_function", "ig_exstack_pkg"]


def _proto(L) -> None:
    def f(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    f("bale_hip_init", C.c_int, C.c_int)
    f("bale_hip_last_error", C.c_char_p)
    f("bale_hip_last_timing", None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))
    f("bale_hip_device_info", C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int))
    f("bale_histo", C.c_int, VP, I64T, VP, I64T, C.c_int, C.c_int, C.POINTER(BaleConfig), VP)
    f("bale_gather", C.c_int, VP, VP, I64T, VP, I64T, C.c_int, C.POINTER(BaleConfig), VP)
    f("bale_sync_errors", C.c_int, VP, P64)
    f("bale_hip_release_stream", C.c_int, VP)
    f("bale_hip_set_alc8r", C.c_int, VP)
    f("bale_histo_noalloc", C.c_int, VP, I64T, VP, I64T, C.c_int, C.c_int, C.POINTER(BaleConfig), VP, VP)
    f("bale_hip_last_strategy", C.c_int)
    f("bale_histo_plan", C.c_int, I64T, I64T, C.c_int, C.POINTER(BaleConfig))
    f("bale_gather_plan", C.c_int, I64T, I64T, C.POINTER(BaleConfig))
    f("bale_fill_splitmix", C.c_int, VP, I64T, VP)
    f("bale_fill_neg", C.c_int, VP, I64T, VP)
    f("bale_gen_index", C.c_int, VP, I64T, I64T, C.c_int, C.c_double, I64T, C.c_uint64, VP)
    f("bale_pack", C.c_int, VP, VP, I64T, VP)
    f("bale_probe_copy", C.c_int, VP, VP, I64T, VP)
    f("bale_probe_random_read", C.c_int, VP, I64T, I64T, VP, VP)
    IP = C.POINTER(Instance)
    f("bale_batch_create", VP, IP, I64T, VP)
    f("bale_batch_launch", C.c_int, VP, VP)
    f("bale_batch_launch_direct", C.c_int, VP, VP)
    f("bale_batch_stats", C.c_int, VP, P64, C.c_int)
    f("bale_batch_destroy", None, VP)
    f("bale_batch_launch_plain", C.c_int, IP, I64T, VP)
    f("bale_batch_launch_each", C.c_int, IP, I64T, VP)
    f("bale_tune", C.c_int, IP, I64T, I64T, C.POINTER(C.c_float), C.c_int, VP)
    f("bale_tune_candidates", C.c_int, C.c_int, I64T, C.POINTER(BaleConfig), C.c_int)
    f("bale_randperm", C.c_int, VP, I64T, I64T, VP)
    f("bale_randperm_shuffle", C.c_int, VP, I64T, I64T, VP)
    f("bale_is_perm", C.c_int, VP, I64T, VP)
    for name in ("rand_permp_agi", "rand_permp_conveyor", "rand_permp_agi_opt"):
        f(name, VP, I64T, C.c_int)
    for name in ("rand_permp_exstack", "rand_permp_exstack2"):
        f(name, VP, I64T, C.c_int, I64T)
    f("is_perm", C.c_int, VP, I64T)
    f("rand_permp", VP, I64T, C.c_int)
    f("bale_spmat_transpose", C.c_int, I64T, I64T, I64T, VP, VP, VP, VP, VP)
    f("bale_spmat_permute", C.c_int, I64T, I64T, I64T, VP, VP, VP, VP, VP, VP, VP)
    f("bale_spmat_gen_er", C.c_int, I64T, C.c_double, C.c_int, C.c_int, C.c_uint64, VP, VP, I64T, P64, VP)
    f("bale_spmat_triangular", C.c_int, I64T, VP, VP, C.c_int, P64, P64, VP)
    SP = C.POINTER(SparseMat)
    for name in ("transpose_matrix_agi", "transpose_matrix_conveyor", "transpose_matrix"):
        f(name, SP, SP)
    for name in ("transpose_matrix_exstack", "transpose_matrix_exstack2"):
        f(name, SP, SP, I64T)
    for name in ("permute_matrix_agi", "permute_matrix_conveyor", "permute_matrix"):
        f(name, SP, SP, VP, VP)
    for name in ("permute_matrix_exstack", "permute_matrix_exstack2"):
        f(name, SP, SP, VP, VP, I64T)
    f("gen_erdos_renyi_graph_dist", SP, C.c_int, C.c_double, I64T, I64T, I64T)
    f("init_matrix", SP, I64T, I64T, I64T)
    f("clear_matrix", None, SP)
    f("compare_matrix", C.c_int, SP, SP)
    f("bale_toposort", C.c_int, I64T, VP, VP, VP, VP, VP, VP, P64, VP)
    for name in ("toposort_matrix_agi", "toposort_matrix_exstack", "toposort_matrix_exstack2",
                 "toposort_matrix_convey", "toposort_matrix_oo", "toposort_matrix_cooler",
                 "toposort_matrix_exstack_orig"):
        f(name, C.c_double, VP, VP, SP, SP)
    f("is_upper_triangular", C.c_int, SP, I64T)
    f("is_lower_triangular", C.c_int, SP, I64T)
    f("bale_all_alloc", VP, I64T, I64T)
    f("bale_all_free", None, VP)
    f("bale_memput", C.c_int, VP, VP, I64T)
    f("bale_memget", C.c_int, VP, VP, I64T)
    f("bale_convey_new", VP, I64T, I64T, C.c_size_t)
    f("bale_convey_begin", C.c_int, VP, C.c_size_t)
    f("bale_convey_push", C.c_int, VP, VP, VP, I64T, VP)
    f("bale_convey_advance", C.c_int, VP, C.c_int, VP)
    f("bale_convey_pull", C.c_int, VP, I64T, C.POINTER(VP), P64, P64)
    f("bale_convey_unpull", C.c_int, VP, I64T, I64T)
    f("bale_convey_reset", C.c_int, VP)
    f("bale_convey_free", C.c_int, VP)
    f("bale_memcpy_d2d", C.c_int, VP, VP, I64T, VP)
    f("histo_agi", C.c_double, VP, I64T, VP)
    f("histo_conveyor", C.c_double, VP, I64T, VP)
    for name in DROPIN_HISTO_BUF:
        f(name, C.c_double, VP, I64T, VP, I64T)
    f("ig_agi", C.c_double, VP, VP, I64T, VP)
    f("ig_conveyor", C.c_double, VP, VP, I64T, VP)
    for name in DROPIN_IG_BUF:
        f(name, C.c_double, VP, VP, I64T, VP, I64T)


def lib():
    """The loaded libbale_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -m bale_old_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        _proto(L)
        _lib = L
    return _lib


def last_error() -> str:
    return lib().bale_hip_last_error().decode()


def check(code: int) -> int:
    if code != OK:
        raise BaleError(code, last_error())
    return code


def _ptr(t) -> int | None:
    if t is None:
        return None
    return t.data_ptr() if hasattr(t, "data_ptr") else int(t)


def _stream(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream if isinstance(stream, int) else stream.cuda_stream


def device_info() -> tuple[int, int]:
    cus, lds = C.c_int(), C.c_int()
    check(lib().bale_hip_device_info(C.byref(cus), C.byref(lds)))
    return cus.value, lds.value


def histo(index, counts, *, nbins=None, packed=False, cfg: BaleConfig | None = None, stream=None,
          sync=True) -> int:
    """counts[index[i]] += 1 on the GPU (torch tensors on cuda)."""
    import torch
    dtype = {torch.int64: I64, torch.int32: I32}[counts.dtype]
    nb = counts.numel() if nbins is None else nbins
    rc = lib().bale_histo(_ptr(index), index.numel(), _ptr(counts), nb, dtype,
                          PACKED if packed else PLAIN, C.byref(cfg) if cfg else None, _stream(stream))
    check(rc)
    return sync_errors(stream) if sync else 0


def gather(tgt, index, table, *, packed=False, cfg: BaleConfig | None = None, stream=None,
           sync=True) -> int:
    """tgt[i] = table[index[i]] on the GPU (torch tensors on cuda)."""
    rc = lib().bale_gather(_ptr(tgt), _ptr(index), index.numel(), _ptr(table), table.numel(),
                           PACKED if packed else PLAIN, C.byref(cfg) if cfg else None, _stream(stream))
    check(rc)
    return sync_errors(stream) if sync else 0


def sync_errors(stream=None) -> int:
    bad = C.c_int64(0)
    check(lib().bale_sync_errors(_stream(stream), C.byref(bad)))
    return bad.value


def histo_plan(n: int, nbins: int, dtype: int = I64) -> BaleConfig:
    c = BaleConfig()
    check(lib().bale_histo_plan(n, nbins, dtype, C.byref(c)))
    return c


def gather_plan(n: int, table_len: int) -> BaleConfig:
    c = BaleConfig()
    check(lib().bale_gather_plan(n, table_len, C.byref(c)))
    return c


def fill_splitmix(table, stream=None) -> None:
    check(lib().bale_fill_splitmix(_ptr(table), table.numel(), _stream(stream)))


def fill_neg(table, stream=None) -> None:
    check(lib().bale_fill_neg(_ptr(table), table.numel(), _stream(stream)))


def gen_index(out, nbins: int, *, dist: int = 0, s: float = 1.0, run_mean: int = 0, seed: int = 1,
              stream=None) -> None:
    """Seeded index stream on the GPU (dist 0 uniform, 1 Zipf(s); run_mean>1 clusters)."""
    check(lib().bale_gen_index(_ptr(out), out.numel(), nbins, dist, float(s), run_mean,
                               seed & 0xFFFFFFFFFFFFFFFF, _stream(stream)))


def pack(pck, index, stream=None) -> None:
    check(lib().bale_pack(_ptr(pck), _ptr(index), index.numel(), _stream(stream)))


def last_timing() -> tuple[float, float, float]:
    g, h, d = C.c_double(), C.c_double(), C.c_double()
    lib().bale_hip_last_timing(C.byref(g), C.byref(h), C.byref(d))
    return g.value, h.value, d.value


def randperm(perm, seed: int, stream=None) -> None:
    """perm (int64 device tensor of length n) = the dart-throwing random
    permutation of 0..n-1 (rand_permp_agi at one PE, spmat/spmat_agi.upc:56)."""
    check(lib().bale_randperm(_ptr(perm), perm.numel(), seed, _stream(stream)))


def randperm_shuffle(perm, seed: int, stream=None) -> None:
    """perm (int64 device tensor of length n) = the Fisher-Yates permutation
    of 0..n-1 (rand_permp_agi_opt at one PE,
    apps/randperm_src/alternates/randperm_agi_opt.upc)."""
    check(lib().bale_randperm_shuffle(_ptr(perm), perm.numel(), seed, _stream(stream)))


def is_perm(perm, stream=None) -> bool:
    """Whether perm (int64 device tensor) holds 0..n-1 once each (on the GPU)."""
    r = lib().bale_is_perm(_ptr(perm), perm.numel(), _stream(stream))
    if r < 0:
        check(r)
    return r == 1


# ---- sparse matrices (CSR as device int64 tensors) -------------------------------------
def spmat_gen_er(n: int, p: float, unit_diag: int = 0, mode: int = 3, seed: int = 1, stream=None):
    """(offset, nonzero) of an Erdos-Renyi matrix generated on the GPU
    (gen_erdos_renyi_graph_dist's method, spmat/spmat_utils.upc:554)."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    nnz = C.c_int64(0)
    check(lib().bale_spmat_gen_er(n, p, unit_diag, mode, seed, _ptr(off), None, 0, C.byref(nnz),
                                  _stream(stream)))
    nz = torch.empty(max(nnz.value, 1), dtype=torch.int64, device=dev)
    check(lib().bale_spmat_gen_er(n, p, unit_diag, mode, seed, _ptr(off), _ptr(nz), nz.numel(),
                                  C.byref(nnz), _stream(stream)))
    return off, nz[: nnz.value]


def spmat_transpose(numcols: int, offset, nonzero, t_offset=None, t_nonzero=None, stream=None):
    """(t_offset, t_nonzero): the transpose (transpose_matrix_agi, spmat/spmat_agi.upc:176)."""
    import torch
    numrows = offset.numel() - 1
    if t_offset is None:
        t_offset = torch.empty(numcols + 1, dtype=torch.int64, device=offset.device)
    if t_nonzero is None:
        t_nonzero = torch.empty(max(nonzero.numel(), 1), dtype=torch.int64, device=offset.device)
    check(lib().bale_spmat_transpose(numrows, numcols, nonzero.numel(), _ptr(offset),
                                     _ptr(nonzero) if nonzero.numel() else None, _ptr(t_offset),
                                     _ptr(t_nonzero), _stream(stream)))
    return t_offset, t_nonzero[: nonzero.numel()]


def spmat_permute(numcols: int, offset, nonzero, rperminv, cperminv, stream=None):
    """(p_offset, p_nonzero): row i to row rperminv[i], column j to cperminv[j]
    (permute_matrix_agi, spmat/spmat_agi.upc:119)."""
    import torch
    numrows = offset.numel() - 1
    p_offset = torch.empty(numrows + 1, dtype=torch.int64, device=offset.device)
    p_nonzero = torch.empty(max(nonzero.numel(), 1), dtype=torch.int64, device=offset.device)
    check(lib().bale_spmat_permute(numrows, numcols, nonzero.numel(), _ptr(offset),
                                   _ptr(nonzero) if nonzero.numel() else None, _ptr(rperminv),
                                   _ptr(cperminv), _ptr(p_offset), _ptr(p_nonzero), _stream(stream)))
    return p_offset, p_nonzero[: nonzero.numel()]


def spmat_triangular(offset, nonzero, upper: bool, stream=None) -> tuple[int, int]:
    """(entries on the wrong side of the diagonal, rows without a diagonal entry)."""
    wrong, nopivot = C.c_int64(0), C.c_int64(0)
    check(lib().bale_spmat_triangular(offset.numel() - 1, _ptr(offset),
                                      _ptr(nonzero) if nonzero.numel() else None, int(upper),
                                      C.byref(wrong), C.byref(nopivot), _stream(stream)))
    return wrong.value, nopivot.value


def toposort(offset, nonzero, t_offset, t_nonzero, stream=None):
    """(rperm, cperm, levels) for a morally upper-triangular matrix and its
    transpose (toposort_matrix_agi, apps/topo_src/toposort_agi.upc:53)."""
    import torch
    n = offset.numel() - 1
    rperm = torch.empty(max(n, 1), dtype=torch.int64, device=offset.device)
    cperm = torch.empty(max(n, 1), dtype=torch.int64, device=offset.device)
    lev = C.c_int64(0)
    check(lib().bale_toposort(n, _ptr(offset), _ptr(nonzero) if nonzero.numel() else None,
                              _ptr(t_offset), _ptr(t_nonzero) if t_nonzero.numel() else None,
                              _ptr(rperm), _ptr(cperm), C.byref(lev), _stream(stream)))
    return rperm[:n], cperm[:n], lev.value

