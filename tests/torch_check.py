"""Independent device-side checkers for the full-size parity tests.

splitmix64 restated with torch int64 arithmetic (wrapping multiply and add,
logical shifts built from arithmetic ones), so a 2^28-element gather can be
checked on the GPU without a host round trip.  It is pinned against the
oracle and the published splitmix64 outputs in tests/test_full_size.py.
"""
from __future__ import annotations

import torch

M64 = (1 << 64) - 1


def _s64(x: int) -> int:
    """Two's-complement int64 value of a 64-bit pattern."""
    x &= M64
    return x - (1 << 64) if x >> 63 else x


GOLDEN = _s64(0x9E3779B97F4A7C15)
C1 = _s64(0xBF58476D1CE4E5B9)
C2 = _s64(0x94D049BB133111EB)


def _shr(z: torch.Tensor, k: int) -> torch.Tensor:
    """Logical right shift of int64 bit patterns."""
    return (z >> k) & ((1 << (64 - k)) - 1)


def splitmix64(j: torch.Tensor) -> torch.Tensor:
    """First splitmix64 output with state j (table[j] of the gather workload)."""
    z = j + GOLDEN
    z = (z ^ _shr(z, 30)) * C1
    z = (z ^ _shr(z, 27)) * C2
    return z ^ _shr(z, 31)


def gather_matches_splitmix(tgt: torch.Tensor, idx: torch.Tensor, chunk: int = 1 << 24) -> bool:
    """tgt[i] == splitmix64(idx[i]) for every i, checked chunk by chunk."""
    n = idx.numel()
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        if not torch.equal(tgt[a:b], splitmix64(idx[a:b])):
            return False
    return True
