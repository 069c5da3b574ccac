"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares with the reference's prototypes, and rejects bad arguments on
the host before touching a device."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import bale_old_amd as B

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "bale_hip.h"

# the reference's hot-path prototypes (apps/histo_src/histo.h:47-50,
# histo_alternates.h:43-46, apps/ig_src/ig.h:49-52, ig_alternates.h:42-45)
REFERENCE_SYMBOLS = [
    "histo_agi", "histo_exstack", "histo_exstack2", "histo_conveyor",
    "histo_exstack2_goto", "histo_exstack2_cyclic", "histo_exstack_function", "histo_exstack2_function",
    "ig_agi", "ig_exstack", "ig_exstack2", "ig_conveyor",
    "ig_exstack2_cyclic", "ig_exstack2_goto", "ig_exstack_function", "ig_exstack_pkg",
]


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # a name followed by "(" is a function, unless "(*" opens a function-pointer field
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\((?!\s*\*)", text)) - {"defined", "sizeof"})


def test_library_loads_and_exports_every_declared_symbol():
    L = B.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", str(B.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    decl = declared_functions()
    assert set(REFERENCE_SYMBOLS) <= set(decl)
    missing = [f for f in decl if f not in exported]
    assert not missing, missing
    for f in decl:
        getattr(L, f)


def test_reference_prototypes_are_kept():
    text = HEADER.read_text()
    assert re.search(r"double histo_agi\(int64_t \*index, int64_t len, int64_t \*counts\);", text)
    assert re.search(r"double ig_agi\(int64_t \*tgt, int64_t \*index, int64_t l_num_req, int64_t \*table\);", text)
    assert re.search(r"double histo_exstack\(int64_t \*pckindx, int64_t len, int64_t \*lcounts, int64_t buf_cnt\);", text)
    assert re.search(r"double ig_conveyor\(int64_t \*tgt, int64_t \*pckindx, int64_t l_num_req, int64_t \*ltable\);", text)


def test_header_compiles_as_c():
    src = '#include "bale_hip.h"\nint main(void){return (int)sizeof(bale_config);}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"), "-x", "c",
                        "-", "-fsyntax-only"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr

def test_struct_layouts_match_ctypes(tmp_path):
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "bale_hip.h"\n'
           'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(bale_config), sizeof(bale_instance),'
           ' offsetof(bale_instance, index), offsetof(bale_instance, cfg));return 0;}\n')
    (tmp_path / "l.c").write_text(src)
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(tmp_path / "l.c"), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert got == [C.sizeof(B.BaleConfig), C.sizeof(B.Instance), B.Instance.index.offset,
                   B.Instance.cfg.offset]


def test_host_side_argument_checks():
    L = B.lib()
    assert L.bale_histo(None, -1, None, 10, B.I64, B.PLAIN, None, None) == B.ERR_ARG
    assert "negative" in B.last_error()
    assert L.bale_histo(None, 0, None, 10, B.I64, B.PLAIN, None, None) == B.OK  # empty: nothing to do
    assert L.bale_histo(None, 5, None, 10, B.I64, B.PLAIN, None, None) == B.ERR_ARG
    assert L.bale_histo(1, 5, 1, 0, B.I64, B.PLAIN, None, None) == B.ERR_ARG
    assert L.bale_histo(1, 5, 1, 10, 7, B.PLAIN, None, None) == B.ERR_ARG
    assert L.bale_histo(1, 5, 1, 10, B.I64, 9, None, None) == B.ERR_ARG
    assert L.bale_gather(None, None, -3, None, 10, B.PLAIN, None, None) == B.ERR_ARG
    assert L.bale_gather(1, 1, 5, 1, 0, B.PLAIN, None, None) == B.ERR_ARG
    assert L.bale_gen_index(1, 5, 10, 3, 1.0, 0, 1, None) == B.ERR_ARG
    assert L.bale_gen_index(1, 5, 10, 1, -1.0, 0, 1, None) == B.ERR_ARG
    assert L.bale_probe_copy(1, 1, 15, None) == B.ERR_ARG


@pytest.mark.parametrize("field,value,words", [
    ("strategy", 7, "strategy"), ("block", 300, "block"), ("unroll", 5, "unroll"),
    ("tile_log2", 40, "tile_log2"), ("merge", 3, "merge"), ("groups_per_cu", 99, "groups_per_cu"),
    ("hot", 3, "hot"), ("task_log2", 5, "task_log2"), ("task_log2", 31, "task_log2"),
])
def test_configs_are_checked_on_the_host(field, value, words):
    """A configuration no shape can take is refused with BALE_ERR_ARG and a
    reason before any device is touched (the pointers are never read)."""
    L = B.lib()
    cfg = B.config(B.BUCKET)
    setattr(cfg, field, value)
    assert L.bale_histo(1, 5, 1, 10, B.I64, B.PLAIN, C.byref(cfg), None) == B.ERR_ARG
    assert words in B.last_error()
    cfg = B.config(B.DIRECT)
    setattr(cfg, field, value)
    if field != "hot" or value != 2:
        assert L.bale_gather(1, 1, 5, 1, 10, B.PLAIN, C.byref(cfg), None) == B.ERR_ARG


def test_hot2_and_lds_are_histogram_or_bucket_only():
    L = B.lib()
    cfg = B.config(B.DIRECT, hot=2)
    assert L.bale_histo(1, 5, 1, 10, B.I64, B.PLAIN, C.byref(cfg), None) == B.ERR_ARG
    cfg = B.config(B.LDS)
    assert L.bale_gather(1, 1, 5, 1, 10, B.PLAIN, C.byref(cfg), None) == B.ERR_ARG
    assert "gather strategy" in B.last_error()


def test_adjacent_kernels_check_arguments_on_the_host():
    """randperm, transpose, permute, generator, toposort: sizes and modes are
    refused before any device is touched (the pointers are never read)."""
    L = B.lib()
    assert L.bale_randperm(None, -1, 1, None) == B.ERR_ARG
    assert L.bale_randperm(None, 0, 1, None) == B.OK
    assert L.bale_randperm(1, 1 << 30, 1, None) == B.ERR_ARG and "2^30" in B.last_error()
    assert L.bale_is_perm(None, 5, None) == B.ERR_ARG
    assert L.rand_permp_agi(-1, 1) is None and "outside" in B.last_error()
    assert L.rand_permp(1 << 30, 1) is None
    big = 1 << 31
    assert L.bale_spmat_transpose(-1, 4, 0, 1, 1, 1, 1, None) == B.ERR_ARG
    assert L.bale_spmat_transpose(4, 4, big, 1, 1, 1, 1, None) == B.ERR_ARG and "2^31" in B.last_error()
    assert L.bale_spmat_transpose(1 << 32, 4, 8, 1, 1, 1, 1, None) == B.ERR_ARG
    assert L.bale_spmat_transpose(4, 4, 8, None, 1, 1, 1, None) == B.ERR_ARG
    assert L.bale_spmat_permute(4, 4, 8, 1, 1, None, 1, 1, 1, None) == B.ERR_ARG
    assert L.bale_spmat_permute(4, -4, 8, 1, 1, 1, 1, 1, 1, None) == B.ERR_ARG
    nnz = C.c_int64(0)
    assert L.bale_spmat_gen_er(10, 0.1, 0, 0, 1, 1, None, 0, C.byref(nnz), None) == B.ERR_ARG  # mode 0
    assert L.bale_spmat_gen_er(10, 1.5, 0, 3, 1, 1, None, 0, C.byref(nnz), None) == B.ERR_ARG
    assert L.bale_spmat_gen_er(-1, 0.1, 0, 3, 1, 1, None, 0, C.byref(nnz), None) == B.ERR_ARG
    assert not L.gen_erdos_renyi_graph_dist(10, 0.1, 0, 0, 1)  # mode 0 is not provided
    w, npiv = C.c_int64(0), C.c_int64(0)
    assert L.bale_spmat_triangular(-1, 1, 1, 1, C.byref(w), C.byref(npiv), None) == B.ERR_ARG
    assert L.bale_toposort(big, 1, 1, 1, 1, 1, 1, None, None) == B.ERR_ARG
    assert L.bale_toposort(5, None, 1, 1, 1, 1, 1, None, None) == B.ERR_ARG
    assert not L.transpose_matrix_agi(None) and "bad sparse matrix" in B.last_error()
    m = B.SparseMat(0, 3, 3, 4, 0, 0, 1, 1, 0, 0)  # 3 x 4, no nonzeros: not square
    assert L.toposort_matrix_agi(1, 1, C.byref(m), C.byref(m)) == -1.0
    assert "square" in B.last_error()


def test_dropin_rejects_negative_length():
    L = B.lib()
    assert L.histo_agi(None, -1, None) == -1.0
    assert L.ig_agi(None, None, -1, None) == -1.0


def test_no_device_is_reported():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert B.lib().bale_hip_init(0) == B.ERR_NODEV
