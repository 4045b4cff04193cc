"""The CPU oracle, pinned before it is trusted.

* bit-exact against the compiled reference's own output (golden vectors
  written by tests/golden/make_golden.py from blackscholes.c);
* the A&S 26.2.17 error bound against the exact normal CDF (a wrong or
  transposed coefficient breaks it);
* Hull's textbook answers, put-call parity, no-arbitrage bounds;
* the closed-form Greeks against finite differences of the closed-form
  price, and the A&S finite-difference Greeks against the closed forms.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.special import ndtr

from conftest import load_golden
from p3arsec_amd import books


def test_golden_bit_exact(oracle):
    for tag in ("config1_4096", "edge"):
        book, ref = load_golden(tag)
        mine, _, _ = oracle.price_f32(book, nthreads=1)
        assert mine.dtype == np.float32
        assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32)), tag


def test_golden_degenerate_inputs(oracle):
    """T = 0, sigma = 0, zero or infinite S and K, NaN inputs...: the oracle
    gives the compiled reference's output for every option: NaN where it
    prints NaN, and otherwise the same "%.18f" text, infinities included (the
    fixture holds the reference's printed values, so a price below 5e-19
    reads back as 0)."""
    book, ref = load_golden("degenerate")
    mine, _, _ = oracle.price_f32(book, nthreads=1)
    nan = np.isnan(ref)
    assert 0 < nan.sum() < book.n and np.isinf(ref).sum() > 0 and np.isfinite(ref).sum() > 20
    assert np.array_equal(np.isnan(mine), nan)
    assert ["%.18f" % x for x in mine[~nan]] == ["%.18f" % x for x in ref[~nan]]


def test_cndf_error_bound_against_exact_phi(oracle):
    x = np.linspace(-8.0, 8.0, 200001)
    err64 = np.abs(oracle.cndf_f64(x) - ndtr(x))
    assert err64.max() <= 7.5e-8          # A&S 26.2.17: |eps(x)| < 7.5e-8
    assert err64.max() >= 1e-8            # and it is the A&S polynomial, not Phi
    err32 = np.abs(oracle.cndf_f32(x.astype(np.float32)).astype(np.float64)
                   - ndtr(x.astype(np.float32).astype(np.float64)))
    assert err32.max() <= 7.5e-8 + 2.5e-7  # plus float rounding of the steps


def test_cndf_special_values(oracle):
    # float rounding of the steps leaves CNDF(0) one ulp below 0.5
    c0 = float(oracle.cndf_f32(np.array([0.0], np.float32))[0])
    assert abs(c0 - 0.5) <= float(np.spacing(np.float32(0.5)))
    # A&S at 0: 1 - 0.3989422804014327 * 1.253314136 = 0.50000000052481
    assert abs(float(oracle.cndf_f64(np.array([0.0]))[0]) - 0.50000000052481) < 1e-13
    x = np.linspace(1e-3, 6, 1001)
    s = oracle.cndf_f64(x) + oracle.cndf_f64(-x)
    assert np.max(np.abs(s - 1.0)) < 1e-15
    y = oracle.cndf_f64(np.linspace(-10, 10, 20001))
    assert np.all(np.diff(y) >= 0)         # monotone
    assert y[0] < 1e-20 and y[-1] == 1.0


@pytest.mark.parametrize("args,call,put", [
    ((42.0, 40.0, 0.1, 0.2, 0.5), 4.7594, 0.8086),     # Hull, worked example
    ((100.0, 100.0, 0.05, 0.3, 1.0), 14.2312, 9.3542),
])
def test_hull_known_answers(oracle, args, call, put):
    L = oracle.lib()
    for fn in (L.bso_price_f32, L.bso_price_f64, L.bso_price_exact):
        assert abs(fn(*args, 0) - call) < 1e-4
        assert abs(fn(*args, 1) - put) < 1e-4


def test_put_call_parity_and_bounds(oracle):
    book = books.make_book(20000, books.SEED_BASE + 11)
    b64 = book.astype(np.float64)
    calls = oracle.price_f64(books.Book(*b64.arrays()[:5], np.zeros(book.n, np.uint8)))
    puts = oracle.price_f64(books.Book(*b64.arrays()[:5], np.ones(book.n, np.uint8)))
    s, k, r, t = b64.s, b64.k, b64.r, b64.t
    fwd = s - k * np.exp(-r * t)
    assert np.max(np.abs((calls - puts) - fwd) / (s + k)) < 1e-12
    eps = 1e-6 * (s + k)
    assert np.all(calls >= np.maximum(fwd, 0) - eps) and np.all(calls <= s + eps)
    assert np.all(puts >= np.maximum(-fwd, 0) - eps) and np.all(puts <= k + eps)


def test_as_price_against_exact_closed_form(oracle):
    book = books.make_book(50000, books.SEED_BASE + 12)
    as64 = oracle.price_f64(book)
    exact = oracle.price_exact(book)
    s, k = book.s.astype(np.float64), book.k.astype(np.float64)
    # each of the two CNDF values is within 7.5e-8 of Phi
    assert np.all(np.abs(as64 - exact) <= 7.5e-8 * (s + k) + 1e-12)
    f32, _, _ = oracle.price_f32(book, nthreads=4)
    err = np.abs(f32.astype(np.float64) - as64)
    assert err.max() < 1e-4                 # the reference's ERR_CHK threshold


def test_closed_form_greeks_against_fd_of_closed_form(oracle):
    book = books.make_book(3000, books.SEED_BASE + 13)
    g = oracle.greeks_exact(book)
    L = oracle.lib()
    h = 1e-5
    for i in range(0, book.n, 37):
        s, k, r, v, t = (float(a[i]) for a in book.arrays()[:5])
        ty = int(book.otype[i])
        P = lambda s_, v_, r_, t_: L.bso_price_exact(s_, k, r_, v_, t_, ty)  # noqa: E731
        fd = [
            (P(s * (1 + h), v, r, t) - P(s * (1 - h), v, r, t)) / (2 * h * s),
            (P(s, v * (1 + h), r, t) - P(s, v * (1 - h), r, t)) / (2 * h * v),
            (P(s, v, r * (1 + h), t) - P(s, v, r * (1 - h), t)) / (2 * h * r),
            -(P(s, v, r, t * (1 + h)) - P(s, v, r, t * (1 - h))) / (2 * h * t),
        ]
        scale = [1.0, s * np.sqrt(t), k * t, s + k]
        for j in range(4):
            assert abs(fd[j] - g[1 + j, i]) <= 1e-5 * scale[j], (i, j, fd[j], g[1 + j, i])
        assert abs(g[0, i] - P(s, v, r, t)) < 1e-12 * (s + k)


def test_fd_of_as_price_close_to_closed_form(oracle):
    # The A&S price's derivatives differ from the exact Greeks by the
    # derivative of the approximation error (|eps'| ~ 5e-7), amplified by
    # 1/(sigma sqrt(T)) on the d1 path: up to ~6e-5 of the scale below.
    book = books.make_book(20000, books.SEED_BASE + 14)
    fd = oracle.greeks_fd(book, 1e-3)
    ex = oracle.greeks_exact(book)
    s, k, t = (book.s.astype(np.float64), book.k.astype(np.float64),
               book.t.astype(np.float64))
    scale = [s + k, np.ones_like(s), s * np.sqrt(t), k * t, s + k]
    for j in range(5):
        assert np.max(np.abs(fd[j] - ex[j]) / scale[j]) < 2e-4, j


def test_static_blocks_thread_invariance(oracle):
    book = books.make_book(100003, books.SEED_BASE + 15)
    one, _, n1 = oracle.price_f32(book, nthreads=1)
    many, _, n7 = oracle.price_f32(book, nthreads=7, passes=3)
    assert n1 == 1 and n7 == 7
    assert np.array_equal(one.view(np.uint32), many.view(np.uint32))


def test_parsec_text_roundtrip(tmp_path):
    book = books.make_book(257, books.SEED_BASE + 16)
    p = tmp_path / "in.txt"
    books.write_parsec_input(str(p), book)
    lines = p.read_text().splitlines()
    assert int(lines[0]) == 257 and len(lines) == 258
    f = lines[1].split()
    assert len(f) == 9 and f[6] in ("P", "C")
    assert np.float32(float(f[0])) == book.s[0]


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/clang"), reason="no clang for the ThreadSanitizer build")
def test_static_blocks_under_thread_sanitizer():
    """The static-block pool (bso_batch_f32) built with -fsanitize=thread:
    7 threads x 3 passes, bit-identical to one thread, no race reported."""
    native = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
    subprocess.run(["make", "-s", "-C", native, "oracle_pool_tsan"], check=True)
    p = subprocess.run([os.path.join(native, "oracle_pool_tsan")], capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66"))
    assert p.returncode == 0 and "ThreadSanitizer" not in p.stderr, p.stderr[-4000:]


def test_c_input_writer_matches_python_writer(tmp_path):
    """pyoracle.write_parsec_input (C) writes the same text as
    books.write_parsec_input, ragged values and both option types included."""
    import filecmp
    import pyoracle
    book = books.make_book(5003, books.SEED_BASE + 17)
    a, b = tmp_path / "py.txt", tmp_path / "c.txt"
    books.write_parsec_input(str(a), book)
    pyoracle.write_parsec_input(str(b), book)
    assert filecmp.cmp(a, b, shallow=False)


@pytest.mark.skipif(not os.access(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                                               "oracle", "_ref", "blackscholes_ff"), os.X_OK),
                    reason="compiled reference not built (needs the reference sources)")
def test_reference_cpu_leg_times_the_compiled_reference(oracle):
    """bench.py's reference leg: runs oracle/_ref/blackscholes_ff on a small
    book, finds its ROI from its output, and its prices equal the port's."""
    import sys
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
    if root not in sys.path:
        sys.path.insert(0, root)
    import bench
    book = books.make_book(20_000, books.SEED_BASE + 18)
    port, _, _ = oracle.price_f32(book, nthreads=2)
    r = bench.reference_cpu(book, 2, port)
    assert r["kind"] == "reference" and r["cores"] == 2 and r["value"] > 0
    assert r["outputs_equal_to_port"]
