"""Concurrent callers on the MI355X: the reference's worker threads call the
pricing functions concurrently (BS:315-348, :607-658), so the C ABI must be
reentrant.  Eight host threads (ctypes releases the GIL during each call)
price their own books on their own streams, and call the scalar forms, in a
loop; every result must equal the single-threaded one bit for bit.  Also the
timing-only events that bracket passes on two streams."""
import ctypes
import threading

import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu


def test_eight_threads_on_their_own_streams(gpu):
    L = gpu.load()
    nthreads, rounds = 8, 20
    bks = [books.make_book(50_000 + 997 * i, books.SEED_BASE + 130 + i) for i in range(nthreads)]
    dbs = [gpu.DeviceBook(b) for b in bks]
    want = [gpu.price(db).download(np.float32, b.n) for db, b in zip(dbs, bks)]
    want_greeks = [[o.download(np.float32, b.n) for o in gpu.greeks(db)] for db, b in zip(dbs, bks)]
    probe = [(float(b.s[7]), float(b.k[7]), float(b.r[7]), float(b.v[7]), float(b.t[7]), int(b.otype[7]))
             for b in bks]
    errors = []

    def worker(i):
        try:
            st = gpu.Stream()
            out = gpu.alloc_out(bks[i].n)
            gouts = [gpu.alloc_out(bks[i].n) for _ in range(5)]
            sc = ctypes.c_float()
            for r in range(rounds):
                if r % 2 == 0:
                    rc = L.bs_price_f32(*dbs[i].ptrs(), out.ptr, bks[i].n, st.handle)
                else:
                    rc = L.bs_price_greeks_f32(*dbs[i].ptrs(), *[o.ptr for o in gouts], bks[i].n,
                                               st.handle)
                assert rc == 0, L.bs_last_error()
                s, k, rr, v, t, ty = probe[i]
                assert L.bs_price_scalar_f32(s, k, rr, v, t, ty, 0.0, ctypes.byref(sc)) == 0
                assert np.float32(sc.value).view(np.uint32) == want[i][7].view(np.uint32)
                assert L.bs_stream_sync(st.handle) == 0
                if r % 2 == 0:
                    got = out.download(np.float32, bks[i].n)
                    assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (i, r)
                else:
                    for o, w in zip(gouts, want_greeks[i]):
                        assert np.array_equal(o.download(np.float32, bks[i].n).view(np.uint32),
                                              w.view(np.uint32)), (i, r)
            for o in [out] + gouts:
                o.free()
            st.destroy()
        except Exception as e:   # reported by the main thread
            errors.append((i, repr(e)))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join(120)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors
    for db in dbs:
        db.free()


def test_timing_events_bracket_passes_on_two_streams(gpu):
    """bs_event_create_timing events (no system-scope fence), as bench.py
    brackets its timed regions: a region of 8 passes alternating over two
    streams, joined at both ends, measures a positive time no shorter than
    one pass timed alone; the outputs read after a device synchronize are the
    map's own bits."""
    L = gpu.load()
    n = 2_000_000
    book = books.make_book(n, books.SEED_BASE + 150)
    dbs = [gpu.DeviceBook(book) for _ in range(2)]
    outs = [gpu.alloc_out(n) for _ in range(2)]
    ref_out = gpu.price(dbs[0])
    want = ref_out.download(np.float32, n)
    st = [gpu.Stream(), gpu.Stream()]
    one0, one1 = gpu.Event(timing=True), gpu.Event(timing=True)
    one0.record(st[0].handle)
    assert L.bs_price_f32(*dbs[0].ptrs(), outs[0].ptr, n, st[0].handle) == 0
    one1.record(st[0].handle)
    one1.sync()
    single = one0.elapsed_ms(one1)
    e0, e1, pre, join = (gpu.Event(timing=True) for _ in range(4))
    pre.record(st[1].handle)
    st[0].wait(pre)
    e0.record(st[0].handle)
    st[1].wait(e0)
    for i in range(8):
        assert L.bs_price_f32(*dbs[i % 2].ptrs(), outs[i % 2].ptr, n, st[i % 2].handle) == 0
    join.record(st[0].handle)
    st[1].wait(join)
    e1.record(st[1].handle)
    e1.sync()
    region = e0.elapsed_ms(e1)
    gpu.call("bs_device_sync")
    print(f"\none pass {single * 1e3:.1f} us, 8 passes on two streams {region * 1e3:.1f} us")
    assert 0.0 < single <= region
    for o in outs:
        assert np.array_equal(o.download(np.float32, n).view(np.uint32), want.view(np.uint32))
    for ev in (one0, one1, e0, e1, pre, join):
        ev.destroy()
    for x in dbs + outs + [ref_out]:
        x.free()
    for s in st:
        s.destroy()
