"""exstack and exstack2 from Python: torch tensors in, torch tensors out.

Mirrors the reference's calls (exstack/exstack.h:94-107, :201-210) in bulk,
as apps/histo_src/histo_exstack.upc:64-76 uses them:

    ex = Exstack(n_pe=8, buf_cnt=1024, pkg_size=8)
    i = 0
    while ex.proceed(done=(i == n)):
        i += ex.push(items[i:], to_pes[i:])           # the prefix that fits the round
        ex.exchange()
        for pe in range(8):
            got, frm = ex.pop(pe)                      # (count, pkg_size) bytes, source PEs
    ex.clear()

There is no CPU path: every call runs on the GPU through libbale_hip.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import BaleError, check, last_error, lib, _stream

VP = C.c_void_p
I64 = C.c_int64


def _proto(L) -> None:
    if getattr(L, "_exstack_proto", False):
        return
    def f(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)
    for k in ("", "2"):
        f(f"bale_exstack{k}_init", VP, I64, I64, C.c_size_t)
        f(f"bale_exstack{k}_push", I64, VP, VP, VP, VP, I64, VP)
        f(f"bale_exstack{k}_pop", I64, VP, I64, C.POINTER(VP), C.POINTER(VP))
        f(f"bale_exstack{k}_reset", C.c_int, VP)
        f(f"bale_exstack{k}_clear", None, VP)
    f("bale_exstack_proceed", I64, VP, C.c_int)
    f("bale_exstack_exchange", C.c_int, VP, VP)
    f("bale_exstack_pop_thread", I64, VP, I64, I64, C.POINTER(VP))
    f("bale_exstack2_proceed", I64, VP, C.c_int, VP)
    L._exstack_proto = True


def _rc(code: int) -> int:
    if code < 0:
        raise BaleError(code, last_error())
    return code


def _bytes(addr: int, count: int, width: int, stream) -> torch.Tensor:
    out = torch.empty((count, width), dtype=torch.uint8, device="cuda")
    if count:
        check(lib().bale_memcpy_d2d(out.data_ptr(), addr, count * width, _stream(stream)))
    return out


class Exstack:
    """Bulk-synchronous aggregation buffers for `n_pe` virtual PEs (kind 1),
    or the asynchronous exstack2 form (kind 2)."""

    def __init__(self, n_pe: int, buf_cnt: int, pkg_size: int = 8, kind: int = 1):
        L = lib()
        _proto(L)
        self.k = "" if kind == 1 else "2"
        h = getattr(L, f"bale_exstack{self.k}_init")(n_pe, buf_cnt, pkg_size)
        if not h:
            raise BaleError(-1, last_error())
        self.h = C.c_void_p(h)
        self.n_pe, self.buf_cnt, self.pkg = n_pe, buf_cnt, pkg_size

    def push(self, items: torch.Tensor, to_pes: torch.Tensor, from_pes: torch.Tensor | None = None,
             stream=None) -> int:
        """Push the longest prefix that fits the round; returns its length."""
        n = to_pes.numel()
        for t in (to_pes, from_pes):
            if t is not None and (t.dtype != torch.int64 or not t.is_cuda or not t.is_contiguous()):
                raise ValueError("PE numbers must be contiguous cuda int64 tensors")
        if from_pes is not None and from_pes.numel() != n:
            raise ValueError("from_pes and to_pes differ in length")
        if not items.is_cuda or not items.is_contiguous() or items.numel() * items.element_size() != n * self.pkg:
            raise ValueError(f"items must be a contiguous cuda tensor of {n} x {self.pkg} bytes")
        fp = from_pes.data_ptr() if from_pes is not None else None
        return _rc(getattr(lib(), f"bale_exstack{self.k}_push")(self.h, items.data_ptr(), fp,
                                                                 to_pes.data_ptr(), n, _stream(stream)))

    def proceed(self, done: bool, stream=None) -> bool:
        if self.k == "":
            return _rc(lib().bale_exstack_proceed(self.h, int(done))) != 0
        return _rc(lib().bale_exstack2_proceed(self.h, int(done), _stream(stream))) != 0

    def exchange(self, stream=None) -> None:
        _rc(lib().bale_exstack_exchange(self.h, _stream(stream)))

    def pop(self, pe: int, stream=None) -> tuple[torch.Tensor, torch.Tensor]:
        """What pe received this round, grouped by source: ((count, pkg) uint8, (count,) int64 sources)."""
        p, s = C.c_void_p(), C.c_void_p()
        cnt = _rc(getattr(lib(), f"bale_exstack{self.k}_pop")(self.h, pe, C.byref(p), C.byref(s)))
        frm = _bytes(s.value or 0, cnt, 8, stream).view(torch.int64).reshape(cnt)
        return _bytes(p.value or 0, cnt, self.pkg, stream), frm

    def pop_thread(self, pe: int, from_pe: int, stream=None) -> torch.Tensor:
        """What from_pe sent pe this round, in push order ((count, pkg) uint8)."""
        p = C.c_void_p()
        cnt = _rc(lib().bale_exstack_pop_thread(self.h, pe, from_pe, C.byref(p)))
        return _bytes(p.value or 0, cnt, self.pkg, stream)

    def reset(self) -> None:
        _rc(getattr(lib(), f"bale_exstack{self.k}_reset")(self.h))

    def clear(self) -> None:
        if self.h:
            getattr(lib(), f"bale_exstack{self.k}_clear")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass
