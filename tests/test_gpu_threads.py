"""Concurrent callers on the MI355X: the reference's worker threads call the
pricing functions concurrently (BS:315-348, :607-658), so the C ABI must be
reentrant.  Eight host threads (ctypes releases the GIL during each call)
price their own books on their own streams, and call the scalar forms, in a
loop; every result must equal the single-threaded one bit for bit."""
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
