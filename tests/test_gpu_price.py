"""Parity of the HIP pricing kernels with the oracle (runs on the MI355X).

Bar: every FP32 price within 1e-4 of the reference's own output (the
reference's ERR_CHK threshold, blackscholes.c:333-340); the fptype=double map
within 1e-9 of the FP64 restatement; the launch shape (alignment, grid cap,
unroll, non-temporal) never changes a bit.  Every test calls through the C ABI.
"""
import os
import ctypes

import numpy as np
import pytest

from conftest import load_golden
from p3arsec_amd import books

pytestmark = pytest.mark.gpu
TOL = 1e-4
# bs_price_fast_f32 (hardware transcendentals) is documented to exceed the
# reference's check on a handful of options per 10M; it is held to 2e-4.
TOL_FAST = 2e-4


def run(capi, book, mode="ref", **kw):
    dt = np.float64 if mode == "f64" else np.float32
    tdt = np.int32 if mode == "i32" else np.uint8
    db = capi.DeviceBook(book, dtype=dt, type_dtype=tdt, pad=kw.pop("pad", 0))
    out_pad = kw.pop("out_pad", 0)
    out = capi.price(db, mode=mode, out_pad=out_pad)
    got = out.download(dt, book.n, offset_bytes=out_pad * np.dtype(dt).itemsize)
    db.free()
    out.free()
    return got


def reset_config(capi):
    capi.set_config(capi.CFG_MAX_BLOCKS, 0)
    capi.set_config(capi.CFG_VECTOR_WIDTH, 0)
    capi.set_config(capi.CFG_NONTEMPORAL, 1)
    capi.set_config(capi.CFG_PIPELINE, -1)
    capi.set_config(capi.CFG_BLOCK_THREADS, 0)


def err_stats(got, ref):
    e = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    return float(e.max()) if e.size else 0.0, float(e.mean()) if e.size else 0.0


@pytest.mark.parametrize("tag", ["config1_4096", "edge"])
@pytest.mark.parametrize("mode", ["ref", "i32", "fast"])
def test_golden_vectors(gpu, oracle, tag, mode):
    book, ref = load_golden(tag)
    got = run(gpu, book, mode)
    mx, mean = err_stats(got, ref)
    exact = oracle.price_exact(book)
    mx64, mean64 = err_stats(got, exact)
    same = float(np.mean(got.view(np.uint32) == ref.view(np.uint32)))
    print(f"\n{tag} {mode}: vs reference max {mx:.3e} mean {mean:.3e} "
          f"bit-identical {same:.4f}; vs FP64 closed form max {mx64:.3e} mean {mean64:.3e}")
    assert mx <= (TOL_FAST if mode == "fast" else TOL)


@pytest.mark.parametrize("mode", ["ref", "i32"])
def test_degenerate_inputs_match_reference(gpu, mode):
    """Options outside the distributions (T = 0, sigma = 0, S or K zero or
    infinite, 0/0, negative T, NaN inputs, magnitudes near 1e30 and 1e-30):
    NaN exactly where the reference gives NaN, the same infinities, and every
    finite price within 1e-4 of the reference's or, for prices above 1e2,
    within 1e-6 relative (eight float ulps)."""
    book, ref = load_golden("degenerate")
    got = run(gpu, book, mode)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    inf = np.isinf(ref)
    assert np.array_equal(got[inf], ref[inf])
    fin = ~nan & ~inf
    d = np.abs(got[fin].astype(np.float64) - ref[fin].astype(np.float64))
    assert np.all(d <= np.maximum(1e-4, 1e-6 * np.abs(ref[fin].astype(np.float64))))


def test_degenerate_inputs_fp64(gpu, oracle):
    """fptype = double on the same degenerate book: NaN where the FP64
    restatement gives NaN, and the same infinities and finite values."""
    book, _ = load_golden("degenerate")
    got = run(gpu, book.astype(np.float64), "f64")
    ref = oracle.price_f64(book)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    fin = ok & np.isfinite(ref)
    assert np.array_equal(got[ok & ~fin], ref[ok & ~fin])
    assert np.all(np.abs(got[fin] - ref[fin]) <= 1e-9 * np.maximum(1.0, np.abs(ref[fin])))


def test_random_book_1m(gpu, oracle):
    book = books.make_book(1 << 20, books.SEED_BASE + 21)
    ref, _, _ = oracle.price_f32(book)
    for mode in ("ref", "fast"):
        got = run(gpu, book, mode)
        mx, mean = err_stats(got, ref)
        print(f"\n1M {mode}: max {mx:.3e} mean {mean:.3e} "
              f"bit-identical {np.mean(got == ref):.4f}")
        assert mx <= (TOL_FAST if mode == "fast" else TOL)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 255, 256, 257, 1023, 1025, 4099])
def test_ragged_sizes(gpu, oracle, n):
    book = books.make_book(n, books.SEED_BASE + 22 + n)
    ref, _, _ = oracle.price_f32(book)
    got = run(gpu, book)
    assert err_stats(got, ref)[0] <= TOL


def test_empty_book(gpu):
    L = gpu.load()
    assert L.bs_price_f32(*([None] * 7), 0, None) == 0


def test_launch_shape_and_alignment_do_not_change_bits(gpu):
    book = books.make_book(100_003, books.SEED_BASE + 23)
    base = run(gpu, book)
    variants = [run(gpu, book, pad=1), run(gpu, book, out_pad=1),   # scalar path
                run(gpu, book, "i32")]
    try:
        for key, val in ((gpu.CFG_MAX_BLOCKS, 7), (gpu.CFG_VECTOR_WIDTH, 2),
                         (gpu.CFG_VECTOR_WIDTH, 4), (gpu.CFG_NONTEMPORAL, 0),
                         (gpu.CFG_PIPELINE, 1), (gpu.CFG_PIPELINE, 0),
                         (gpu.CFG_BLOCK_THREADS, 64), (gpu.CFG_BLOCK_THREADS, 128)):
            gpu.set_config(key, val)
            variants.append(run(gpu, book))
            reset_config(gpu)
    finally:
        reset_config(gpu)
    for v in variants:
        assert np.array_equal(v.view(np.uint32), base.view(np.uint32))


def test_full_size_10m_against_oracle_and_parity(gpu, oracle):
    n = 10_000_000
    book = books.make_book(n, books.SEED_BASE + 2)
    got = run(gpu, book)
    ref, _, _ = oracle.price_f32(book)
    mx, mean = err_stats(got, ref)
    print(f"\n10M vs oracle: max {mx:.3e} mean {mean:.3e} "
          f"bit-identical {np.mean(got == ref):.5f}")
    ex = err_stats(got, oracle.price_exact(book))
    r64 = err_stats(got, oracle.price_f64(book.astype(np.float64)))
    print(f"10M vs FP64 closed form: max {ex[0]:.3e} mean {ex[1]:.3e}; "
          f"vs FP64 restatement: max {r64[0]:.3e} mean {r64[1]:.3e}")
    assert ex[0] <= 2e-4 and r64[0] <= 2e-4
    fast = run(gpu, book, "fast")
    fmx, fmean = err_stats(fast, ref)
    print(f"10M fast vs oracle: max {fmx:.3e} mean {fmean:.3e} "
          f"over 1e-4: {int(np.sum(np.abs(fast.astype(np.float64) - ref) > 1e-4))}")
    assert mx <= TOL and fmx <= TOL_FAST
    # size-independent check: put-call parity across a call book and a put book
    calls = run(gpu, books.Book(*book.arrays()[:5], np.zeros(n, np.uint8)))
    puts = run(gpu, books.Book(*book.arrays()[:5], np.ones(n, np.uint8)))
    s, k, r, t = (a.astype(np.float64) for a in (book.s, book.k, book.r, book.t))
    fwd = s - k * np.exp(-r * t)
    assert np.max(np.abs(calls - puts - fwd)) < 2 * TOL


@pytest.mark.parametrize("n,seed", [(1 << 20, 3), (10_000_000, 2)])
def test_fp64_map(gpu, oracle, n, seed):
    """fptype = double: against the FP64 restatement of the reference and the
    exact closed form; the FP32 map's distance from it (printed) stays inside
    the reference's 1e-4.  10,000,000 options is the FP64 configuration's book
    (the benchmark's seed)."""
    book = books.make_book(n, books.SEED_BASE + seed)
    got = run(gpu, book, "f64")
    ref = oracle.price_f64(book.astype(np.float64))
    mx, mean = err_stats(got, ref)
    f32 = run(gpu, book)
    d32 = err_stats(f32, got)
    ex = err_stats(got, oracle.price_exact(book))
    print(f"\n{n} options: FP64 vs FP64 oracle max {mx:.3e} mean {mean:.3e}; vs closed form max "
          f"{ex[0]:.3e} mean {ex[1]:.3e}; FP32 - FP64 max {d32[0]:.3e} mean {d32[1]:.3e}")
    assert mx <= 1e-9
    assert d32[0] < TOL
    s, k = book.s.astype(np.float64), book.k.astype(np.float64)
    assert np.all(np.abs(got - oracle.price_exact(book)) <= 7.5e-8 * (s + k) + 1e-9)


def test_cndf_batch(gpu, oracle):
    x = np.linspace(-9, 9, 100_003)
    db = gpu.DeviceBuffer.from_array(x.astype(np.float32))
    out = gpu.DeviceBuffer(x.size * 4)
    gpu.call("bs_cndf_f32", db.ptr, out.ptr, x.size, None)
    got = out.download(np.float32, x.size)
    ref = oracle.cndf_f32(x.astype(np.float32))
    assert np.max(np.abs(got.astype(np.float64) - ref)) < 3e-7
    # special values: the reference's CNDF gives 1 at +inf, 0 at -inf, NaN at NaN
    sp = np.array([np.inf, -np.inf, np.nan, 1e30, -1e30, 25.0, -25.0], np.float32)
    ds = gpu.DeviceBuffer.from_array(sp)
    os_ = gpu.DeviceBuffer(sp.size * 4)
    gpu.call("bs_cndf_f32", ds.ptr, os_.ptr, sp.size, None)
    gpu.call("bs_stream_sync", None)
    gs, rs = os_.download(np.float32, sp.size), oracle.cndf_f32(sp)
    assert np.array_equal(np.isnan(gs), np.isnan(rs)) and np.isnan(gs).sum() == 1
    assert np.array_equal(gs[~np.isnan(gs)], rs[~np.isnan(rs)].astype(np.float32))
    d64 = gpu.DeviceBuffer.from_array(x)
    o64 = gpu.DeviceBuffer(x.size * 8)
    gpu.call("bs_cndf_f64", d64.ptr, o64.ptr, x.size, None)
    assert np.max(np.abs(o64.download(np.float64, x.size) - oracle.cndf_f64(x))) < 1e-15


def test_scalar_entry_points(gpu, oracle):
    L = gpu.load()
    out = ctypes.c_float()
    gpu.check(L.bs_price_scalar_f32(42.0, 40.0, 0.1, 0.2, 0.5, 0, 0.0, ctypes.byref(out)), "s")
    assert abs(out.value - 4.7594) < 1e-4
    gpu.check(L.bs_price_scalar_f32(42.0, 40.0, 0.1, 0.2, 0.5, 1, 0.0, ctypes.byref(out)), "s")
    assert abs(out.value - 0.8086) < 1e-4
    gpu.check(L.bs_cndf_scalar_f32(0.0, ctypes.byref(out)), "c")
    assert abs(out.value - 0.5) < 1e-6


def test_host_pointer_form_and_thread_compat(gpu, oracle):
    book = books.make_book(10_007, books.SEED_BASE + 24)
    ref, _, _ = oracle.price_f32(book)
    arrs = [np.ascontiguousarray(a) for a in (book.s, book.k, book.r, book.v, book.t)]
    ty = book.otype.astype(np.int32)
    out = np.zeros(book.n, np.float32)
    L = gpu.load()
    gpu.check(L.bs_price_host_f32(*[a.ctypes.data for a in arrs], ty.ctypes.data,
                                  out.ctypes.data, book.n, 3, 0), "host")
    assert err_stats(out, ref)[0] <= TOL
    # bs_thread drop-in: 3 threads, static blocks of N/3, the N%3 tail untouched
    prices = np.full(book.n, -1.0, np.float32)
    gpu.check(L.bs_bind_globals(book.n, 3, *[a.ctypes.data for a in arrs], ty.ctypes.data,
                                prices.ctypes.data), "bind")
    for tid in range(3):
        t = ctypes.c_int(tid)
        gpu.check(L.bs_thread_compat(ctypes.byref(t)), "compat")
    blk = book.n // 3
    assert err_stats(prices[:3 * blk], ref[:3 * blk])[0] <= TOL
    assert np.all(prices[3 * blk:] == -1.0)


@pytest.mark.parametrize("n,offset", [(5000, 0), (1, 0), (255, 0), (256, 0), (257, 0),
                                      (300_001, 0), (5000, 4), (2048, 4)])
def test_repack_aos(gpu, n, offset):
    """Every field of every record lands in its array, bit for bit, for books
    that fill the kernel's 256-record steps, end inside one, or are smaller
    than one; offset = 4 places the records 4 bytes past a 16-byte boundary
    (the reference's records are only 4-byte aligned)."""
    book = books.make_book(n, books.SEED_BASE + 25)
    rec = np.dtype([("s", "<f4"), ("strike", "<f4"), ("r", "<f4"), ("divq", "<f4"),
                    ("v", "<f4"), ("t", "<f4"), ("OptionType", "S1"), ("pad", "V3"),
                    ("divs", "<f4"), ("DGrefval", "<f4")])
    assert rec.itemsize == 36
    aos = np.zeros(n, rec)
    for f, a in (("s", book.s), ("strike", book.k), ("r", book.r), ("v", book.v),
                 ("t", book.t)):
        aos[f] = a
    chars = np.array([b"P", b"C", b"X", b"p"])
    pick = np.arange(n) % 4
    aos["OptionType"] = chars[pick]
    aos["divq"] = 7.0
    raw = np.zeros(offset + aos.nbytes, np.uint8)
    raw[offset:] = aos.view(np.uint8)
    daos = gpu.DeviceBuffer.from_array(raw)
    outs = [gpu.DeviceBuffer(n * 4) for _ in range(5)]
    dty = gpu.DeviceBuffer(n)
    gpu.call("bs_repack_aos_f32", daos.ptr + offset, n, *[o.ptr for o in outs], dty.ptr, None)
    gpu.call("bs_stream_sync", None)
    for o, a in zip(outs, (book.s, book.k, book.r, book.v, book.t)):
        assert np.array_equal(o.download(np.float32, n).view(np.uint32), a.view(np.uint32))
    assert np.array_equal(dty.download(np.uint8, n), (pick == 0).astype(np.uint8))
    for b in [daos, dty] + outs:
        b.free()


def test_invalid_pointers_rejected_on_device(gpu):
    L = gpu.load()
    small = gpu.DeviceBuffer(64)
    big = gpu.DeviceBuffer(4096)
    host = np.zeros(1024, np.float32)
    p = [big.ptr] * 5 + [big.ptr, big.ptr]
    # allocation shorter than n elements
    assert L.bs_price_f32(*p[:4], small.ptr, big.ptr, big.ptr, 1000, None) == -1
    # unpinned host memory
    assert L.bs_price_f32(*p[:6], host.ctypes.data, 1000, None) == -1
    gpu.call("bs_price_f32", *p, 1000, None)
    gpu.call("bs_stream_sync", None)


def test_fp64_primitives_against_the_c_library(gpu):
    """The FP64 policy's log64 and recip64 on the GPU against the host C
    library over ~8.7M log and ~8.4M reciprocal inputs (every binade,
    subnormals, [0.5, 2], 1 +- k ulp, special values): log64 within 2 ulp,
    recip64 within 1 ulp, special values identical (tests/native/fp64_prims.hip)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "fp64_prims")
    assert os.path.exists(exe), "build it: make -C tests/native fp64_prims"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print("\n" + p.stdout.strip())
    assert p.returncode == 0, p.stdout + p.stderr
