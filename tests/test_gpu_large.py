"""Books past 2^31 options on the MI355X (64-bit indexing), through the C ABI.

A random 2^20-option book is tiled across arrays of 2^31 + 3 options with
device-to-device copies (about 54 GB for the price map, 97 GB with the
Greeks outputs); every tile must then price bit-identically to the small book.
The tiles checked are the first, one in the middle, the last below 2^31 and
the three options above it (the tail the kernel prices one per work-item).
Skipped when the device cannot hold the arrays.
"""
import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu

N_BIG = (1 << 31) + 3
TILE = 1 << 20
CHECK_TILES = [0, 1024, 2047, 2048]


def alloc(capi, nbytes, held):
    L = capi.load()
    import ctypes
    p = ctypes.c_void_p()
    rc = L.bs_malloc(ctypes.byref(p), nbytes)
    if rc != 0:   # the caller's finally frees what is held
        pytest.skip(f"cannot allocate {nbytes / 2**30:.1f} GiB on the device")
    held.append(p.value)
    return p.value


def tile_book(capi, db, held):
    """Big arrays s, k, r, v, t (float) and type (uint8), tiled from db."""
    big = [alloc(capi, N_BIG * 4, held) for _ in range(5)] + [alloc(capi, N_BIG, held)]
    src = db.ptrs()
    for t in range((N_BIG + TILE - 1) // TILE):
        cnt = min(TILE, N_BIG - t * TILE)
        for j in range(6):
            es = 4 if j < 5 else 1
            capi.call("bs_memcpy_d2d", big[j] + t * TILE * es, src[j], cnt * es, None)
    capi.call("bs_stream_sync", None)
    return big


def check_tiles(capi, out_ptr, want):
    for t in CHECK_TILES:
        cnt = min(TILE, N_BIG - t * TILE)
        got = np.empty(cnt, np.float32)
        capi.call("bs_memcpy_d2h", got.ctypes.data, out_ptr + t * TILE * 4, cnt * 4, None)
        capi.call("bs_stream_sync", None)
        assert np.array_equal(got.view(np.uint32), want[:cnt].view(np.uint32)), t


def test_price_and_greeks_past_2_31_options(gpu):
    base = books.make_book(TILE, books.SEED_BASE + 120)
    db = gpu.DeviceBook(base)
    ref = gpu.price(db).download(np.float32, TILE)
    gref = [o.download(np.float32, TILE) for o in gpu.greeks(db)]
    held = []
    try:
        big = tile_book(gpu, db, held)
        out = alloc(gpu, N_BIG * 4, held)
        gpu.call("bs_price_f32", *big, out, N_BIG, None)
        gpu.call("bs_stream_sync", None)
        check_tiles(gpu, out, ref)
        gouts = [out] + [alloc(gpu, N_BIG * 4, held) for _ in range(4)]
        gpu.call("bs_memset", out, 0, N_BIG * 4, None)
        gpu.call("bs_price_greeks_f32", *big, *gouts, N_BIG, None)
        gpu.call("bs_stream_sync", None)
        for o, w in zip(gouts, gref):
            check_tiles(gpu, o, w)
    finally:
        for b in held:
            gpu.call("bs_free", b)
        db.free()
