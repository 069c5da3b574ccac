"""The allocator hook (bale_hip_set_alc8r, the reference's convey_alc8r_t of
convey/convey_alc8r.h:53-60): every device allocation the library makes for
itself goes through the installed allocator, every pointer goes back to the
allocator that made it, and an allocator that returns NULL makes each call
fail with BALE_ERR_ALLOC (the drop-ins with -1.0 or NULL) and leak nothing.

The allocators here hand out hipMalloc memory and keep a ledger of what is
live, as the reference's LIFO-checking allocator (convey/alltoallv.c:81-117)
checks its frees.  Failures are forced from the host; every kernel that runs
gets valid memory.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import bale_old_amd as B
from bale_old_amd import mix as MX
from bale_old_amd.convey import Conveyor
from bale_old_amd.exstack import Exstack

pytestmark = pytest.mark.gpu

GRAB = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t, C.c_char_p, C.c_uint64)
FREE = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)


class Alc8rT(C.Structure):
    _fields_ = [("alc8r", C.c_void_p), ("grab", GRAB), ("free", FREE)]


_HIP = None
_KEEP = []  # the library may hold pointers of an allocator until its stream is released


def hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so.7")  # the runtime already loaded in this process
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipMalloc.restype = C.c_int
        _HIP.hipFree.argtypes = [C.c_void_p]
        _HIP.hipFree.restype = C.c_int
    return _HIP


class Ledger:
    """hipMalloc memory with every live pointer recorded; grabs past
    `fail_after` successful ones return NULL."""

    def __init__(self, fail_after=None):
        self.fail_after = fail_after
        self.live = {}
        self.grabbed = 0
        self.refused = 0
        self.bad_frees = []
        self.tags = set()
        self.s = Alc8rT(None, GRAB(self._grab), FREE(self._free))
        _KEEP.append(self)

    def _grab(self, ctx, size, tag, value):
        if self.fail_after is not None and self.grabbed >= self.fail_after:
            self.refused += 1
            return None
        p = C.c_void_p()
        if hip().hipMalloc(C.byref(p), size) != 0 or not p.value:
            return None
        self.grabbed += 1
        self.live[p.value] = (size, tag.decode(), value)
        self.tags.add(tag.decode())
        return p.value

    def _free(self, ctx, ptr):
        if ptr not in self.live:
            self.bad_frees.append(ptr)
            return
        del self.live[ptr]
        hip().hipFree(ptr)

    def __enter__(self):
        B.check(B.lib().bale_hip_set_alc8r(C.byref(self.s)))
        return self

    def __exit__(self, *exc):
        B.lib().bale_hip_set_alc8r(None)


def refuse_all():
    return Ledger(fail_after=0)


def small_histo_inputs(dev, n=200_000, nbins=50_000, seed=3):
    idx = torch.empty(n, dtype=torch.int64, device=dev)
    B.gen_index(idx, nbins, seed=seed)
    torch.cuda.synchronize()
    return idx, nbins


def release(st):
    B.check(B.lib().bale_hip_release_stream(C.c_void_p(st.cuda_stream)))


def test_every_library_allocation_goes_through_the_hook(dev, orc):
    L = B.lib()
    st = torch.cuda.Stream()
    idx, nbins = small_histo_inputs(dev)
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    # the default stream's record and workspace exist before the ledger starts
    L.bale_all_free(L.rand_permp_agi(10_000, 7))
    torch.cuda.synchronize()
    with Ledger() as led:
        with torch.cuda.stream(st):
            B.histo(idx, counts, cfg=B.config(B.BUCKET), stream=st)
            table = torch.empty(1 << 20, dtype=torch.int64, device=dev)
            B.fill_splitmix(table, st)
            gi = torch.empty(1 << 20, dtype=torch.int64, device=dev)
            B.gen_index(gi, 1 << 20, seed=4, stream=st)
            tgt = torch.empty_like(gi)
            B.gather(tgt, gi, table, cfg=B.config(B.BUCKET), stream=st)
        st.synchronize()
        B.sync_errors(st)
        after_calls = dict(led.live)
        assert led.grabbed >= 2  # the stream's error record and workspace at least
        p = L.bale_all_alloc(1000, 8)
        assert p and p in led.live
        L.bale_all_free(p)
        assert p not in led.live
        perm = L.rand_permp_agi(10_000, 7)
        assert perm and perm in led.live
        L.bale_all_free(perm)
        cv = Conveyor(8, 4096, 8)
        assert len(led.live) == len(after_calls) + 3  # queue, PE numbers, delivery
        cv.free()
        ex = Exstack(4, 256, 8)
        assert len(led.live) == len(after_calls) + 11
        ex.clear()
        assert led.live == after_calls
    # results are right with the caller's memory
    want = np.zeros(nbins, np.int64)
    orc.histo_agi(idx.cpu().numpy(), want)
    assert np.array_equal(counts.cpu().numpy(), want)
    assert np.array_equal(tgt.cpu().numpy(), table.cpu().numpy()[gi.cpu().numpy()])
    # the stream's own memory goes back to this allocator when it is released,
    # although the default allocator is installed again
    release(st)
    assert led.live == {} and led.bad_frees == []
    assert "capi.hip" in led.tags


def test_an_allocator_that_refuses_makes_calls_fail_cleanly(dev, orc):
    L = B.lib()
    st = torch.cuda.Stream()
    idx, nbins = small_histo_inputs(dev)
    counts = torch.zeros(nbins, dtype=torch.int64, device=dev)
    with refuse_all() as led:
        with pytest.raises(B.BaleError) as e:
            B.histo(idx, counts, cfg=B.config(B.BUCKET), stream=st)
        assert e.value.code == B.ERR_ALLOC
        assert L.bale_all_alloc(10, 8) is None
        assert "allocation" in B.last_error()
        assert L.rand_permp_agi(1000, 1) is None
        host_idx = idx[:1000].cpu().numpy().copy()
        host_counts = np.zeros(nbins, np.int64)
        assert L.histo_agi(host_idx.ctypes.data, 1000, host_counts.ctypes.data) == -1.0
        assert "staging" in B.last_error()
        assert L.bale_convey_new(8, 1024, 8) is None
        with pytest.raises(B.BaleError):
            Exstack(4, 64, 8)
        assert led.refused >= 6 and led.grabbed == 0
    # nothing was counted, and with the default allocator back every call works
    assert int(counts.sum()) == 0
    with torch.cuda.stream(st):
        B.histo(idx, counts, cfg=B.config(B.BUCKET), stream=st)
    st.synchronize()
    want = np.zeros(nbins, np.int64)
    orc.histo_agi(idx.cpu().numpy(), want)
    assert np.array_equal(counts.cpu().numpy(), want)
    release(st)


def test_a_failure_at_every_allocation_of_a_batch_leaks_nothing(dev, orc):
    """Batch creation on a stream that already has its record: the k-th of its
    allocations fails for every k; each attempt fails with nothing left
    allocated, and the batch that gets all its memory is exact."""
    L = B.lib()
    st = torch.cuda.Stream()
    params = MX.mix_params(24, 11)
    mix = MX.Mix(params, dev, st)
    # bucketed, LDS, hot and gather instances in one batch
    cfgs = []
    for i, p in enumerate(params):
        if p.kind == 1:
            cfgs.append(B.config(B.DIRECT))
        else:
            cfgs.append(B.config([B.BUCKET, B.LDS, B.DIRECT][i % 3], hot=1 if i % 2 else 0))
    mix.set_configs(cfgs)
    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        B.histo(torch.zeros(1, dtype=torch.int64, device=dev), torch.zeros(1, dtype=torch.int64, device=dev),
                stream=st)  # the stream's record, outside the ledgers
    st.synchronize()
    sp = C.c_void_p(st.cuda_stream)
    with Ledger() as full:
        h = L.bale_batch_create(mix.inst, len(params), sp)
        assert h, B.last_error()
        need = full.grabbed
        L.bale_batch_destroy(h)
        assert full.live == {}
    assert need >= 5
    for k in range(need):
        with Ledger(fail_after=k) as led:
            h = L.bale_batch_create(mix.inst, len(params), sp)
            assert not h, k
            assert led.refused >= 1
            assert led.live == {} and led.bad_frees == [], k
    mix.reset_counts()
    torch.cuda.synchronize()
    with Ledger(fail_after=need) as led:
        h = L.bale_batch_create(mix.inst, len(params), sp)
        assert h
        B.check(L.bale_batch_launch(h, sp))
        st.synchronize()
        B.sync_errors(st)
        L.bale_batch_destroy(h)
        assert led.live == {}
    for p, (tv, iv, _) in zip(params, mix.views):
        if p.kind == 0:
            want = np.zeros(p.table_len, np.int64)
            orc.histo_agi(iv.cpu().numpy(), want)
            assert np.array_equal(tv.cpu().numpy(), want)
    release(st)


def test_allocator_needs_both_functions():
    bad = Alc8rT(None, GRAB(lambda *a: None), FREE())
    assert B.lib().bale_hip_set_alc8r(C.byref(bad)) == B.ERR_ARG
    assert B.lib().bale_hip_set_alc8r(None) == 0
