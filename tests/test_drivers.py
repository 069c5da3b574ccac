"""The C drivers (bale_old_amd/host/*.c): the reference programs' command
line, output and self-checks, run on the GPU through libbale_hip.so."""
import re
import subprocess
from pathlib import Path

import numpy as np
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
    assert "options (defaults in brackets)" in r.stderr and "-M MASK" in r.stderr


def test_randperm_driver_all_models():
    r = run("randperm", "-n", "300000", "-M", "31")
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Running randperm on 1 threads" in r.stderr
    for m in ("AGI", "EXSTACK", "EXSTACK2", "CONVEYOR"):
        assert re.search(rf"rand_permp_{m}:\s+[\d.]+", r.stderr), m
    assert re.search(r"rand_permp_agi_opt :\s+[\d.]+", r.stderr)  # the alternate model
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
    r = run("topo", "-n", "20000", "-M", "31")
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Running toposort on 1 threads" in r.stderr
    assert re.search(r"Input matrix has 20000 rows and \d+ nonzeros", r.stdout)
    assert len(re.findall(r"\s+[\d.]+ seconds", r.stderr)) == 5
    assert "ERROR" not in r.stderr and "ERROR" not in r.stdout



def write_mm(path, n, rows, cols, field="pattern", nnz=None, banner=None):
    lines = [banner or f"%%MatrixMarket matrix coordinate {field} general", "% written by the test",
             f"{n} {n} {len(rows) if nnz is None else nnz}"]
    for r, c in zip(rows, cols):
        lines.append(f"{r + 1} {c + 1}" + (" 1" if field == "integer" else ""))
    path.write_text("\n".join(lines) + "\n")


def permuted_upper_triangle(n, seed):
    """A unit upper-triangular pattern under random row and column
    permutations, entries in random order: what toposort undoes."""
    rng = np.random.default_rng(seed)
    r, c = np.triu_indices(n)
    keep = (r == c) | (rng.random(r.size) < 0.05)
    r, c = r[keep], c[keep]
    rp, cp = rng.permutation(n), rng.permutation(n)
    order = rng.permutation(r.size)
    return rp[r][order], cp[c][order]


@pytest.mark.parametrize("field", ["pattern", "integer"])
def test_topo_driver_reads_matrix_market(tmp_path, field):
    """-f: the matrix comes from a Matrix Market file
    (apps/topo_src/toposort.upc:252, spmat/spmat_utils.upc:217-365) and every
    model toposorts it."""
    n = 700
    rows, cols = permuted_upper_triangle(n, seed=11)
    f = tmp_path / "m.mtx"
    write_mm(f, n, rows, cols, field)
    r = run("topo", "-f", str(f), "-M", "31")
    assert r.returncode == 0, r.stderr + r.stdout
    assert f"Input matrix has {n} rows and {rows.size} nonzeros" in r.stdout
    assert len(re.findall(r"\s+[\d.]+ seconds", r.stderr)) == 5
    assert "ERROR" not in r.stderr and "ERROR" not in r.stdout


@pytest.mark.parametrize("case", ["banner", "array", "count", "range"])
def test_topo_driver_rejects_bad_matrix_market(tmp_path, case):
    n = 50
    rows, cols = permuted_upper_triangle(n, seed=3)
    f = tmp_path / "bad.mtx"
    if case == "banner":
        f.write_text("not a matrix market file\n")
    elif case == "array":
        write_mm(f, n, rows, cols, banner="%%MatrixMarket matrix array real general")
    elif case == "count":
        write_mm(f, n, rows, cols, nnz=rows.size + 1)
    else:
        rows = rows.copy()
        rows[7] = n  # written 1-based as n + 1
        write_mm(f, n, rows, cols)
    r = run("topo", "-f", str(f))
    assert r.returncode != 0
    assert "ERROR: read_matrix_mm" in r.stderr
