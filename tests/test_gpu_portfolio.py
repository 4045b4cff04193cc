"""Portfolio reductions (bs_portfolio_f32) on the MI355X, through the C ABI.

* The sums equal the exact sums (math.fsum) of the map's own per-option
  outputs (bs_price_f32 / bs_price_greeks_f32) times the weights, up to the
  rounding of an FP64 reduction: each term w*p is exact in FP64, so only the
  additions round.
* Ragged and empty books, misaligned arrays (the one-option-per-work-item
  path), weights, and bit-reproducibility of repeated calls.
* Against the reference (the oracle's bit-exact prices), the price sum is
  within sum_i |w_i| * 1e-4 (the per-option parity bound); the observed
  difference is printed.
"""
import math

import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu


def portfolio(capi, dbook, n, greeks, weights=None):
    out = capi.DeviceBuffer(40)
    pad = dbook.pad   # the weights get the book's misalignment too
    wbuf = None
    if weights is not None:
        wbuf = capi.DeviceBuffer.from_array(
            np.concatenate([np.zeros(pad, np.float32), weights.astype(np.float32)]))
    wptr = (wbuf.ptr + 4 * pad) if wbuf is not None else None
    capi.call("bs_portfolio_f32", *dbook.ptrs(), wptr, n, 1 if greeks else 0, out.ptr, None)
    capi.call("bs_stream_sync", None)
    sums = out.download(np.float64, 5 if greeks else 1)
    out.free()
    if wbuf is not None:
        wbuf.free()
    return sums


def map_outputs(capi, dbook, n, greeks):
    if greeks:
        return [o.download(np.float32, n) for o in capi.greeks(dbook)]
    return [capi.price(dbook).download(np.float32, n)]


def exact_sums(cols, w):
    res, scale = [], []
    for c in cols:
        terms = c.astype(np.float64) * (w.astype(np.float32).astype(np.float64) if w is not None else 1.0)
        res.append(math.fsum(terms.tolist()))
        scale.append(float(np.sum(np.abs(terms))))
    return np.array(res), np.array(scale)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 255, 1023, 100_003, 1_000_000])
@pytest.mark.parametrize("greeks", [0, 1])
def test_sums_match_the_map(gpu, n, greeks):
    book = books.make_book(max(n, 1), books.SEED_BASE + 80 + n % 97).slice(0, n) if n else None
    if n == 0:
        db = gpu.DeviceBook(books.make_book(1, books.SEED_BASE + 80))
        sums = portfolio(gpu, db, 0, greeks)
        assert np.all(sums == 0.0)
        db.free()
        return
    rng = np.random.default_rng(n + 7 * greeks)
    w = rng.uniform(-2.0, 5.0, n).astype(np.float32)
    db = gpu.DeviceBook(book)
    cols = map_outputs(gpu, db, n, greeks)
    for weights in (None, w):
        got = portfolio(gpu, db, n, greeks, weights)
        ref, scale = exact_sums(cols, weights)
        assert np.all(np.abs(got - ref) <= 1e-13 * scale + 1e-300), (got, ref)
        again = portfolio(gpu, db, n, greeks, weights)
        assert np.array_equal(got.view(np.uint64), again.view(np.uint64))   # fixed order
    db.free()


def test_misaligned_arrays_take_the_scalar_path(gpu):
    n = 50_001
    book = books.make_book(n, books.SEED_BASE + 90)
    w = np.random.default_rng(3).uniform(0.5, 2.0, n).astype(np.float32)
    aligned = gpu.DeviceBook(book)
    shifted = gpu.DeviceBook(book, pad=1)
    cols = map_outputs(gpu, aligned, n, 1)
    for db in (aligned, shifted):
        got = portfolio(gpu, db, n, True, w)
        ref, scale = exact_sums(cols, w)
        assert np.all(np.abs(got - ref) <= 1e-13 * scale)
    aligned.free()
    shifted.free()


def test_price_sum_against_the_reference(gpu, oracle):
    n = 1_000_000
    book = books.make_book(n, books.SEED_BASE + 91)
    w = np.random.default_rng(4).uniform(0.0, 10.0, n).astype(np.float32)
    ref, _, _ = oracle.price_f32(book)
    db = gpu.DeviceBook(book)
    got = portfolio(gpu, db, n, False, w)[0]
    want = math.fsum((ref * w.astype(np.float64)).tolist())
    bound = float(np.sum(np.abs(w.astype(np.float64)))) * 1e-4
    print(f"\nportfolio 1M: |GPU sum - reference sum| = {abs(got - want):.3e} "
          f"(bound {bound:.3e}, relative {abs(got - want) / abs(want):.2e})")
    assert abs(got - want) <= bound
    db.free()


def test_argument_errors(gpu):
    L = gpu.load()
    db = gpu.DeviceBook(books.make_book(8, books.SEED_BASE + 92))
    assert L.bs_portfolio_f32(*db.ptrs(), None, 8, 0, None, None) == -1          # sums NULL
    out = gpu.DeviceBuffer(40)
    assert L.bs_portfolio_f32(*db.ptrs(), None, -1, 0, out.ptr, None) == -1      # n < 0
    page = gpu.DeviceBuffer(4096)                                                  # 5 sums need 40 B
    assert L.bs_portfolio_f32(*db.ptrs(), None, 8, 1, page.ptr + 4088, None) == -1
    db.free()
    out.free()
    page.free()
