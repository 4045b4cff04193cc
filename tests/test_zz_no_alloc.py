"""bs_portfolio_f32_noalloc -- the portfolio sums with the device's fixed
scratch, allocating nothing -- against bs_portfolio_f32 through the C ABI:
bit-identical sums on two calls in a row, the second with a smaller book (so
the fixed scratch still holds the first call's partials), in both forms (price
only, price and Greeks), with and without weights, on aligned and misaligned
arrays; sizes past its limit are refused."""
import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu


def sums_of(capi, name, dbook, n, greeks, wptr):
    out = capi.DeviceBuffer(40)
    capi.call("bs_memset", out.ptr, 0xff, 40, None)      # NaN bits: every sum must be written
    capi.call(name, *dbook.ptrs(), wptr, n, 1 if greeks else 0, out.ptr, None)
    capi.call("bs_stream_sync", None)
    got = out.download(np.float64, 5 if greeks else 1)
    out.free()
    return got


@pytest.mark.parametrize("pad", [0, 1])
def test_noalloc_equals_allocating_form_twice_in_a_row(gpu, pad):
    for n in (3_000_001, 40_961):          # second call smaller than the first
        book = books.make_book(n, books.SEED_BASE + 300 + n % 7)
        db = gpu.DeviceBook(book, pad=pad)
        w = np.random.default_rng(n).uniform(-1.0, 3.0, n).astype(np.float32)
        wbuf = gpu.DeviceBuffer.from_array(np.concatenate([np.zeros(pad, np.float32), w]))
        for greeks in (0, 1):
            for wptr in (None, wbuf.ptr + 4 * pad):
                want = sums_of(gpu, "bs_portfolio_f32", db, n, greeks, wptr)
                got = sums_of(gpu, "bs_portfolio_f32_noalloc", db, n, greeks, wptr)
                print(f"n={n} pad={pad} greeks={greeks} weights={wptr is not None}: {got}")
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        wbuf.free()
        db.free()


def test_noalloc_refuses_books_past_its_scratch(gpu):
    L = gpu.load()
    fn = L.bs_portfolio_f32_noalloc
    out = gpu.DeviceBuffer(40)
    assert fn(*([None] * 7), (1 << 25) + 1, 0, out.ptr, None) == -1
    assert b"2^25" in L.bs_last_error()
    out.free()
