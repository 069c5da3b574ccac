"""Build recipe for libbale_hip.so (gfx950) and the C drivers.

Run as ``python -m bale_old_amd.build`` or through ``__graft_entry__.build()``.
Objects go to ``bale_old_amd/_build/``; the shared library lands in-tree at
``bale_old_amd/libbale_hip.so`` so it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libbale_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["capi.hip", "capi_spmat.hip", "spmat.hip", "toposort.hip", "gen.hip", "gather.hip", "histo.hip", "probe.hip", "batch.hip", "select.hip",
               "convey.hip", "exstack.hip", "randperm.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    noise = [ln for ln in (r.stderr or "").splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "warning generated" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> list[Path]:
    """C drivers mirroring the reference's programs, under the reference's
    binary names (apps/Makefile.am:39)."""
    out = []
    progs = {"histo": "histo", "ig": "ig", "randperm": "randperm", "transpose_matrix": "transpose_matrix",
             "permute_matrix": "permute_matrix", "topo": "toposort"}
    for prog, stem in progs.items():
        src = HOST / f"{stem}.c"
        if not src.exists():
            continue
        exe = BUILD / prog
        _run(["gcc", "-O2", "-std=gnu11", "-Wall", f"-I{ROOT / 'include'}", str(src),
              "-o", str(exe), f"-L{PKG}", "-lbale_hip", f"-Wl,-rpath,{PKG}",
              "-Wl,-rpath,$ORIGIN/..", "-lm"])
        out.append(exe)
    return out


def build_all() -> None:
    build_lib()
    build_host()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
