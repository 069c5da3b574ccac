"""Host-side AddressSanitizer and UndefinedBehaviorSanitizer runs (CPU only).

Four programs, each built with -fsanitize=address,undefined and
-fno-sanitize-recover=all, so any memory error or undefined behaviour ends
the run with a report:
  - tests/sanitize/oracle_check.c: every oracle function on small and edge
    inputs, checked against properties that fix its result (LeakSanitizer on);
  - tests/sanitize/mm_read_check.c: the drivers' Matrix Market reader on good,
    malformed and hostile files (LeakSanitizer on);
  - the six C drivers' option parsing and help text;
  - tests/sanitize/capi_args.c: every C-ABI entry point with bad and good
    arguments, over a build of libbale_hip.so whose host code is sanitized
    (device code as usual), on a machine without a GPU, where each call must
    stop at its argument check or at BALE_ERR_NODEV.

The sanitized library is built into bale_old_amd/_build/asan/ (about a
minute) and reused while the sources are unchanged.
"""
import hashlib
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "tests" / "sanitize"
OUT = ROOT / "bale_old_amd" / "_build" / "asan"
CSRC = ROOT / "bale_old_amd" / "csrc"
HOST = ROOT / "bale_old_amd" / "host"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang"
SAN_FLAGS = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer"]

pytestmark = [
    pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed"),
    pytest.mark.skipif(Path("/dev/kfd").exists(),
                       reason="host-only sanitizer runs are for machines without a GPU"),
]


def run(cmd, env=None, timeout=600, check=True, cwd=None):
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=timeout,
                       env=env, cwd=cwd)
    if check and r.returncode != 0:
        raise AssertionError(f"{' '.join(map(str, cmd))} exited {r.returncode}\n"
                             f"{r.stdout[-4000:]}\n{r.stderr[-8000:]}")
    return r


def san_env(leaks: bool) -> dict:
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = f"detect_leaks={1 if leaks else 0}:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def clang_rt_dir() -> Path:
    dirs = sorted(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux"))
    if not dirs or not Path(HIPCC).exists():
        pytest.skip("ROCm clang and its sanitizer runtimes are not installed")
    return dirs[-1]


def asan_library() -> Path:
    """libbale_hip_asan.so: the library with its host code sanitized."""
    rt = clang_rt_dir()
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.h")) + [ROOT / "include" / "bale_hip.h"]
    h = hashlib.sha256()
    for f in srcs:
        h.update(f.name.encode() + f.read_bytes())
    stamp = OUT / "stamp"
    lib = OUT / "libbale_hip_asan.so"
    if lib.exists() and stamp.exists() and stamp.read_text() == h.hexdigest():
        return lib
    OUT.mkdir(parents=True, exist_ok=True)
    flags = ["--offload-arch=gfx950", "-O1", "-g", "-fPIC", "-std=c++17", "-w",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{ROOT / 'include'}",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]

    def compile_one(src: Path) -> Path:
        obj = OUT / (src.stem + ".o")
        run([HIPCC, *flags, "-c", src, "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, sorted(CSRC.glob("*.hip"))))
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-shared-libasan",
         "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
         "-o", lib, *objs, "-lpthread", f"-Wl,-rpath,{rt}"])
    stamp.write_text(h.hexdigest())
    return lib


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_check"
    run(["gcc", *SAN_FLAGS, "-std=gnu11", "-ffp-contract=off", f"-I{ROOT / 'oracle'}",
         SAN / "oracle_check.c", ROOT / "oracle" / "bale_oracle.c", "-o", exe, "-lm", "-lpthread"])
    r = run([exe], env=san_env(leaks=True))
    assert "0 failed" in r.stderr


def test_matrix_market_reader_under_sanitizers(tmp_path):
    exe = tmp_path / "mm_read_check"
    run(["gcc", *SAN_FLAGS, "-std=gnu11", "-Wall", f"-I{HOST}", SAN / "mm_read_check.c", "-o", exe])
    files = tmp_path / "files"
    files.mkdir()
    r = run([exe, files], env=san_env(leaks=True))
    assert "0 failed" in r.stderr


@pytest.mark.parametrize("prog,stem", [("histo", "histo"), ("ig", "ig"), ("randperm", "randperm"),
                                       ("transpose_matrix", "transpose_matrix"),
                                       ("permute_matrix", "permute_matrix"), ("topo", "toposort")])
def test_driver_options_under_sanitizers(tmp_path, prog, stem):
    lib = ROOT / "bale_old_amd" / "libbale_hip.so"
    if not lib.exists():
        pytest.skip("libbale_hip.so is not built")
    exe = tmp_path / prog
    run(["gcc", *SAN_FLAGS, "-std=gnu11", "-Wall", f"-I{ROOT / 'include'}", HOST / f"{stem}.c",
         "-o", exe, f"-L{lib.parent}", "-lbale_hip", f"-Wl,-rpath,{lib.parent}", "-lm"])
    env = san_env(leaks=False)
    r = run([exe, "-h"], env=env)
    assert "options (defaults in brackets)" in r.stderr
    # options parsed, then the GPU is missing: a clean exit 1
    r = run([exe, "-n", "1000", "-M", "1"], env=env, check=False)
    assert r.returncode == 1, r.stderr[-4000:]
    assert "no GPU" in r.stderr and "Sanitizer" not in r.stderr


def test_capi_argument_paths_under_sanitizers(tmp_path):
    lib = asan_library()
    rt = clang_rt_dir()
    exe = tmp_path / "capi_args"
    run([CLANG, *SAN_FLAGS, "-shared-libasan", f"-I{ROOT / 'include'}", SAN / "capi_args.c",
         "-o", exe, f"-L{lib.parent}", "-lbale_hip_asan", f"-Wl,-rpath,{lib.parent}",
         f"-Wl,-rpath,{rt}", "-lm"])
    r = run([exe], env=san_env(leaks=False), timeout=300)
    assert " 0 failed" in r.stderr, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr
