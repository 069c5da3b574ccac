"""The C drivers (bale_old_amd/host/*.c): the reference programs' command
line, output and self-checks, run on the GPU through libbale_hip.so."""
import re
import subprocess
from pathlib import Path

import pytest

BUILD = Path(__file__).resolve().parent.parent / "bale_old_amd" / "_build"

pytestmark = pytest.mark.gpu


def run(prog, *args):
    exe = BUILD / prog
    if not exe.exists():
        pytest.skip(f"{exe} not built (python -m bale_old_amd.build)")
    return subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("extra", [[], ["-d"], ["-d", "-S", "7"]])
def test_histo_driver_all_models(extra):
    r = run("histo", "-n", "200000", "-T", "5000", "-M", "31", "-c", "1", *extra)
    assert r.returncode == 0, r.stderr
    assert "Running histo on 1 threads" in r.stderr
    # four models plus the four alternates report a time and a rate
    assert len(re.findall(r"seconds\s+[\d.]+ GB/s injection bandwidth", r.stderr)) == 8
    assert "FAILED" not in r.stderr and "ERROR" not in r.stderr


@pytest.mark.parametrize("extra", [[], ["-d"]])
def test_ig_driver_all_models(extra):
    r = run("ig", "-n", "300000", "-T", "70000", "-M", "31", "-c", "1", *extra)
    assert r.returncode == 0, r.stderr
    assert "Running ig on 1 threads" in r.stderr
    assert len(re.findall(r"seconds\s+[\d.]+ GB/s/Node", r.stderr)) == 8
    assert "YOU FAILED" not in r.stderr


def test_histo_driver_help():
    r = run("histo", "-h")
    assert r.returncode == 0
    assert "Usage:" in r.stderr and "-M mask" in r.stderr


def test_randperm_driver_all_models():
    r = run("randperm", "-n", "300000", "-M", "31")
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Running randperm on 1 threads" in r.stderr
    for m in ("AGI", "EXSTACK", "EXSTACK2", "CONVEYOR"):
        assert re.search(rf"rand_permp_{m}:\s+[\d.]+", r.stderr), m
    assert "There is no alternate model here!" in r.stderr
    assert "BALE FAIL" not in r.stderr and "ERROR" not in r.stdout


def test_transpose_driver_all_models():
    r = run("transpose_matrix", "-n", "50000", "-M", "15")
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Running transpose_matrix on 1 threads" in r.stderr
    assert len(re.findall(r"(AGI|Exstack|Exstack2|Conveyor):\s+[\d.]+", r.stderr)) == 4
    assert "ERROR" not in r.stderr


def test_permute_driver_all_models():
    r = run("permute_matrix", "-n", "50000", "-Z", "12")
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Running permute_matrix on 1 threads" in r.stderr
    assert len(re.findall(r"permute_matrix_\w+:\s+[\d.]+", r.stderr)) == 4
    assert "ERROR" not in r.stderr


def test_toposort_driver_all_models():
    r = run("toposort", "-n", "20000", "-M", "31")
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Running toposort on 1 threads" in r.stderr
    assert re.search(r"Input matrix has 20000 rows and \d+ nonzeros", r.stdout)
    assert len(re.findall(r"\s+[\d.]+ seconds", r.stderr)) == 5
    assert "ERROR" not in r.stderr and "ERROR" not in r.stdout

