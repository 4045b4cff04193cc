"""bs_price_passes_f32 (the reference's NUM_RUNS loop in one call) on the
MI355X, through the C ABI.

* Output bit-identical to one bs_price_f32 / bs_price_i32type_f32 pass, for
  books on both sides of the graph threshold (2^25 options), 1 to 100 passes,
  both type widths, misaligned arrays.
* Graph replay: the same call replayed with the inputs rewritten between
  replays and eager single passes interleaved; every replay equals the eager
  result for the inputs it saw (the graph holds pointers, not contents).
* More distinct calls than the cache holds (16), then the first ones again.
"""
import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu


def passes_call(capi, db, out, n, passes, stream=None):
    tb = db.type_dtype.itemsize
    p = db.ptrs()
    capi.call("bs_price_passes_f32", *p[:5], p[5], tb, out.ptr, n, passes, stream)


def eager(capi, db, n):
    fn = "bs_price_f32" if db.type_dtype.itemsize == 1 else "bs_price_i32type_f32"
    out = capi.alloc_out(n)
    capi.call(fn, *db.ptrs(), out.ptr, n, None)
    got = out.download(np.float32, n)
    out.free()
    return got


@pytest.mark.parametrize("n,passes", [(1, 100), (4096, 100), (100_003, 7), (10_000_000, 3),
                                      (1 << 25, 2), ((1 << 25) + 1, 2), (4097, 1)])
@pytest.mark.parametrize("type_dtype", [np.uint8, np.int32])
def test_same_bits_as_single_passes(gpu, n, passes, type_dtype):
    book = books.make_book(n, books.SEED_BASE + 100 + n % 89)
    db = gpu.DeviceBook(book, type_dtype=type_dtype)
    want = eager(gpu, db, n)
    out = gpu.alloc_out(n)
    for _ in range(2):                      # capture, then replay
        passes_call(gpu, db, out, n, passes)
        gpu.call("bs_stream_sync", None)
        assert np.array_equal(out.download(np.float32, n).view(np.uint32), want.view(np.uint32))
    db.free()
    out.free()


def test_misaligned_arrays(gpu):
    n = 5000
    book = books.make_book(n, books.SEED_BASE + 110)
    db = gpu.DeviceBook(book, pad=1)
    want = eager(gpu, db, n)
    out = gpu.alloc_out(n)
    passes_call(gpu, db, out, n, 10)
    gpu.call("bs_stream_sync", None)
    assert np.array_equal(out.download(np.float32, n).view(np.uint32), want.view(np.uint32))
    db.free()
    out.free()


def test_replays_follow_the_inputs(gpu):
    """One captured graph replayed 6 times on a stream; between replays the
    spot array is rewritten and an eager pass runs on another output."""
    n = 8192
    st = gpu.Stream()
    base = books.make_book(n, books.SEED_BASE + 111)
    db = gpu.DeviceBook(base)
    out = gpu.alloc_out(n)
    rng = np.random.default_rng(5)
    for rep in range(6):
        s = (base.s * np.float32(rng.uniform(0.8, 1.2))).astype(np.float32)
        gpu.call("bs_memcpy_h2d", db.ptrs()[0], s.ctypes.data, n * 4, None)
        gpu.call("bs_stream_sync", None)
        want = eager(gpu, db, n)                       # eager work between replays
        passes_call(gpu, db, out, n, 100, st.handle)
        gpu.call("bs_stream_sync", st.handle)
        got = out.download(np.float32, n)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), rep
    st.destroy()
    db.free()
    out.free()


def test_cache_eviction(gpu):
    n = 2048
    book = books.make_book(n, books.SEED_BASE + 112)
    db = gpu.DeviceBook(book)
    want = eager(gpu, db, n)
    outs = [gpu.alloc_out(n) for _ in range(20)]      # 20 distinct keys > 16 cached
    for rnd in range(2):
        for i, o in enumerate(outs):
            passes_call(gpu, db, o, n, 2 + i % 3)
        gpu.call("bs_stream_sync", None)
        for o in outs:
            assert np.array_equal(o.download(np.float32, n).view(np.uint32), want.view(np.uint32))
            gpu.call("bs_memset", o.ptr, 0, n * 4, None)
        gpu.call("bs_stream_sync", None)
    db.free()
    for o in outs:
        o.free()


def test_argument_errors(gpu):
    L = gpu.load()
    book = books.make_book(16, books.SEED_BASE + 113)
    db = gpu.DeviceBook(book)
    out = gpu.alloc_out(16)
    p = db.ptrs()
    assert L.bs_price_passes_f32(*p[:5], p[5], 1, out.ptr, 16, 0, None) == -1    # passes < 1
    assert L.bs_price_passes_f32(*p[:5], p[5], 2, out.ptr, 16, 3, None) == -1    # type width
    assert L.bs_price_passes_f32(*p[:5], p[5], 1, out.ptr, -1, 3, None) == -1    # n < 0
    assert L.bs_price_passes_f32(*p[:5], p[5], 1, out.ptr, 0, 3, None) == 0      # nothing to do
    db.free()
    out.free()
