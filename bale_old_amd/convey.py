"""The device conveyor from Python: torch tensors in, torch tensors out.

Mirrors the reference's conveyor calls (convey/convey.h:184-338) in bulk:

    c = Conveyor(n_pe=8, capacity=1 << 20, item_bytes=8)
    c.begin()
    c.push(items, pes)                 # items: (n, item_bytes) bytes or (n,) int64/int32 ...
    while c.advance(done=True) != DONE:
        for pe in range(8):
            got = c.pull(pe)           # the items sent to pe, in push order
    c.free()

There is no CPU path: every call runs on the GPU through libbale_hip.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import BaleError, check, last_error, lib, _stream

DONE = 0
FAIL = 0
OK = 1


def _rc(code: int) -> int:
    if code < 0:
        raise BaleError(code, last_error())
    return code


class Conveyor:
    """A conveyor delivering fixed-size items to `n_pe` virtual PEs on one GPU."""

    def __init__(self, n_pe: int, capacity: int, item_bytes: int = 8):
        h = lib().bale_convey_new(n_pe, capacity, item_bytes)
        if not h:
            raise BaleError(-1, last_error())
        self.h = C.c_void_p(h)
        self.n_pe, self.capacity, self.item_bytes = n_pe, capacity, item_bytes

    def begin(self, item_bytes: int | None = None) -> int:
        if item_bytes is not None:
            self.item_bytes = item_bytes
        return _rc(lib().bale_convey_begin(self.h, self.item_bytes))

    def push(self, items: torch.Tensor, pes: torch.Tensor, stream=None) -> bool:
        """Queue len(pes) items (a cuda tensor whose rows are item_bytes each).
        False when they do not fit before the next delivery."""
        n = pes.numel()
        if pes.dtype != torch.int64 or not pes.is_cuda or not pes.is_contiguous():
            raise ValueError("pes must be a contiguous cuda int64 tensor")
        if not items.is_cuda or not items.is_contiguous():
            raise ValueError("items must be a contiguous cuda tensor")
        if items.numel() * items.element_size() != n * self.item_bytes:
            raise ValueError(f"items hold {items.numel() * items.element_size()} bytes, "
                             f"expected {n} x {self.item_bytes}")
        return _rc(lib().bale_convey_push(self.h, items.data_ptr(), pes.data_ptr(), n,
                                          _stream(stream))) == OK

    def advance(self, done: bool, stream=None) -> int:
        return _rc(lib().bale_convey_advance(self.h, int(done), _stream(stream)))

    def pull_span(self, pe: int) -> tuple[int, int]:
        """(device address, item count) of what pe has not pulled yet."""
        p, cnt, frm = C.c_void_p(), C.c_int64(), C.c_int64()
        _rc(lib().bale_convey_pull(self.h, pe, C.byref(p), C.byref(cnt), C.byref(frm)))
        return (p.value or 0), cnt.value

    def pull(self, pe: int, stream=None) -> torch.Tensor:
        """The items delivered to pe and not pulled yet, copied into a new
        (count, item_bytes) uint8 cuda tensor."""
        addr, cnt = self.pull_span(pe)
        out = torch.empty((cnt, self.item_bytes), dtype=torch.uint8, device="cuda")
        if cnt:
            check(lib().bale_memcpy_d2d(out.data_ptr(), addr, cnt * self.item_bytes, _stream(stream)))
        return out

    def unpull(self, pe: int, count: int) -> bool:
        return _rc(lib().bale_convey_unpull(self.h, pe, count)) == OK

    def reset(self) -> int:
        return _rc(lib().bale_convey_reset(self.h))

    def free(self) -> None:
        if self.h:
            lib().bale_convey_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
