"""The C++ host tools: the PARSEC `blackscholes` command line and the
compatibility header (include/bs_compat.hpp) driven like the reference's
drivers.  Argument handling runs on any host; pricing needs the GPU."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from p3arsec_amd import books

CLI = os.path.join(ROOT, "p3arsec_amd", "bin", "blackscholes")
COMPAT = os.path.join(ROOT, "tests", "native", "compat_demo")


def test_cli_usage_and_missing_input(tmp_path):
    assert os.path.exists(CLI), "build the tools: make -C p3arsec_amd/csrc tools"
    p = subprocess.run([CLI], capture_output=True, text=True)
    assert p.returncode == 1 and "Usage" in p.stdout
    p = subprocess.run([CLI, "1", str(tmp_path / "none.txt"), str(tmp_path / "o.txt")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "Unable to open" in p.stdout
    bad = tmp_path / "bad.txt"
    bad.write_text("3\n1 2 3\n")
    p = subprocess.run([CLI, "1", str(bad), str(tmp_path / "o.txt")], capture_output=True,
                       text=True)
    assert p.returncode == 1 and "Unable to read" in p.stdout


def write_input_with_ref(path, book, ref):
    tchar = np.where(book.otype != 0, "P", "C")
    with open(path, "w") as f:
        f.write(f"{book.n}\n")
        for i in range(book.n):
            f.write("%.9g %.9g %.9g 0 %.9g %.9g %s 0 %.9g\n" % (
                book.s[i], book.k[i], book.r[i], book.v[i], book.t[i], tchar[i], ref[i]))


@pytest.mark.gpu
def test_cli_reference_workflow(gpu, tmp_path):
    """The reference's own workflow, input text -> output text, on the GPU:
    every price within 1e-4 of the compiled reference (its ERR_CHK count is 0)."""
    book, ref = load_golden("config1_4096")
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    write_input_with_ref(inp, book, ref)
    p = subprocess.run([CLI, "8", str(inp), str(out), "--check"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Num of Options: 4096" in p.stdout and "Num Errors: 0" in p.stdout
    got = books.read_parsec_output(str(out))
    assert np.max(np.abs(got.astype(np.float64) - ref)) <= 1e-4


@pytest.mark.gpu
def test_cli_check_counts_errors_per_pass(gpu, tmp_path):
    """--check with 7 options whose DGrefval is 2e-3 off the reference's price:
    the reference's ERR_CHK counts and reports each once per pass, so 700
    "Error on" lines and "Num Errors: 700" (the other prices are within
    8.4e-5 of the reference, so none of them is flagged)."""
    book, ref = load_golden("config1_4096")
    dg = ref.astype(np.float32).copy()
    bad = [3, 100, 777, 1500, 2048, 3333, 4095]
    for j, i in enumerate(bad):
        dg[i] += np.float32(2e-3 if j % 2 else -2e-3)
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    write_input_with_ref(inp, book, dg)
    p = subprocess.run([CLI, "8", str(inp), str(out), "--check"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("Error on")]
    assert "Num Errors: 700" in p.stdout and len(lines) == 700
    assert [int(ln.split()[2].rstrip(".")) for ln in lines[:7]] == bad


@pytest.mark.gpu
def test_compat_header_forms(gpu, oracle, tmp_path):
    book = books.make_book(1001, books.SEED_BASE + 60)
    ref, _, _ = oracle.price_f32(book)
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    with open(inp, "w") as f:
        f.write(f"{book.n}\n")
        for i in range(book.n):
            f.write("%.9g %.9g %.9g %.9g %.9g %d\n" % (book.s[i], book.k[i], book.r[i],
                                                     book.v[i], book.t[i], book.otype[i]))
    p = subprocess.run([COMPAT, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    cols = np.loadtxt(out, dtype=np.float64)
    for j in (0, 2):                                   # Map run/wait, pointer batch
        assert np.max(np.abs(cols[:, j] - ref)) <= 1e-4
    half = book.n // 2                                 # bs_thread: 2 static blocks, tail dropped
    assert np.max(np.abs(cols[:2 * half, 1] - ref[:2 * half])) <= 1e-4
    assert np.all(cols[2 * half:, 1] == -1.0)
    assert np.max(np.abs(cols[:64, 3] - ref[:64])) <= 1e-4
    x = (np.arange(64) - 32) / 8.0
    assert np.max(np.abs(cols[:64, 4] - oracle.cndf_f32(x.astype(np.float32)))) < 3e-7
    for j in (5, 6):                                   # ParallelFor, SkePU Map
        assert np.max(np.abs(cols[:, j] - ref)) <= 1e-4
        assert np.array_equal(cols[:, j].astype(np.float32), cols[:, 0].astype(np.float32))
    # SkePU lazy sync: 100 map calls upload each input once and download the
    # prices once (at flush); the OpenMP back-end is refused (no host pricing)
    tail = open(out).read().strip().splitlines()[-1].split()
    assert tail[0] == "#" and tail[2] == "1" and tail[4] == "1" and tail[6] == "1", tail


def test_parallel_for_refuses_host_bodies(tmp_path):
    """ParallelFor runs the Black-Scholes body only: any other loop body is a
    compile-time error, so no loop silently runs on the host."""
    src = tmp_path / "bad.cpp"
    src.write_text('#include "bs_compat.hpp"\n'
                   "int main() { bs_compat::ParallelFor pf;\n"
                   "  pf.parallel_for(0, 10, 1, 0, [](long, int) {}); return 0; }\n")
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert p.returncode != 0 and "Black-Scholes body" in p.stderr
