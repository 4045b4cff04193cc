"""The serving loop's host bookkeeping on the CPU (no GPU): the HBM arena
allocator, FIFO admission by memory, batch limits, cancellation, failure
injection and shutdown.  Drives p3arsec_amd/csrc/serve_core.cpp through the
test-only C front in tests/native/serve_core_shim.cpp."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
OK, INVALID, BUDGET, NOT_FOUND, CANCELLED, BUSY, FAILED, SHUTDOWN = 0, -1, -4, -5, -6, -8, -10, -11
QUEUED, RUNNING, DONE, FAILED_ST, CANCELLED_ST = range(5)
PRICE, GREEKS = 1, 2
i64 = ctypes.c_int64


@pytest.fixture(scope="module")
def L():
    subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    # BS_SERVE_CORE_SHIM=libserve_core_shim_asan.so (make -C tests/native asan, run
    # with libasan preloaded) runs the same tests under AddressSanitizer + UBSan.
    lib = ctypes.CDLL(os.path.join(NATIVE, os.environ.get("BS_SERVE_CORE_SHIM",
                                                          "libserve_core_shim.so")))
    vp = ctypes.c_void_p
    P = ctypes.POINTER
    sig = {
        "sc_create": ([i64, i64, i64], vp), "sc_destroy": ([vp], None),
        "sc_register": ([vp, i64, P(i64), P(i64)], ctypes.c_int),
        "sc_release_book": ([vp, i64], ctypes.c_int),
        "sc_submit": ([vp, i64, ctypes.c_int, i64, P(i64)], ctypes.c_int),
        "sc_cancel": ([vp, i64, i64], ctypes.c_int),
        "sc_submit_many": ([vp, P(i64), P(ctypes.c_int), P(i64), i64, i64, P(i64)],
                           ctypes.c_int),
        "sc_enable_completions": ([vp, ctypes.c_int], None),
        "sc_admit_cap": ([vp, i64, i64, P(i64), ctypes.c_int], ctypes.c_int),
        "sc_queued_options": ([vp], i64),
        "sc_create_lanes": ([i64, i64], vp),
        "sc_admit_lane": ([vp, i64, ctypes.c_int, P(i64), ctypes.c_int], ctypes.c_int),
        "sc_queued_lane": ([vp, ctypes.c_int], i64),
        "sc_reap": ([vp, P(i64), P(ctypes.c_int), P(ctypes.c_int), P(i64), i64], i64),
        "sc_admit": ([vp, i64, P(i64), P(i64), ctypes.c_int], ctypes.c_int),
        "sc_complete": ([vp, P(i64), ctypes.c_int, ctypes.c_int, i64], None),
        "sc_fail_queued": ([vp, ctypes.c_int, i64], None),
        "sc_release_result": ([vp, i64], ctypes.c_int),
        "sc_state": ([vp, i64, P(ctypes.c_int), P(ctypes.c_int)], ctypes.c_int),
        "sc_arena_used": ([vp], i64), "sc_arena_peak": ([vp], i64),
        "sc_arena_largest_free": ([vp], i64), "sc_queued": ([vp], i64),
        "sc_stats": ([vp, P(i64)], None),
        "sc_book_bytes": ([i64], i64), "sc_output_bytes": ([i64, ctypes.c_int], i64),
        "sa_create": ([i64], vp), "sa_destroy": ([vp], None),
        "sa_alloc": ([vp, i64, P(i64)], ctypes.c_int), "sa_free": ([vp, i64], ctypes.c_int),
        "sa_used": ([vp], i64), "sa_largest_free": ([vp], i64),
    }
    for name, (a, r) in sig.items():
        f = getattr(lib, name)
        f.argtypes = a
        f.restype = r
    return lib


class Core:
    def __init__(self, L, arena, max_req=0, max_opts=0):
        self.L = L
        self.h = L.sc_create(arena, max_req, max_opts)
        self.clock = 0

    def __del__(self):
        self.L.sc_destroy(self.h)

    def book(self, n):
        i, o = i64(), i64()
        rc = self.L.sc_register(self.h, n, ctypes.byref(i), ctypes.byref(o))
        return rc, i.value, o.value

    def submit(self, book, kind=PRICE):
        t = i64()
        self.clock += 1
        rc = self.L.sc_submit(self.h, book, kind, self.clock, ctypes.byref(t))
        return rc, t.value

    def admit(self, cap=8192):
        t = (i64 * cap)()
        o = (i64 * cap)()
        self.clock += 1
        k = self.L.sc_admit(self.h, self.clock, t, o, cap)
        return list(t[:k]), list(o[:k])

    def complete(self, tickets, ok=True):
        arr = (i64 * max(1, len(tickets)))(*tickets)
        self.clock += 1
        self.L.sc_complete(self.h, arr, len(tickets), 1 if ok else 0, self.clock)

    def state(self, t):
        s, st = ctypes.c_int(), ctypes.c_int()
        rc = self.L.sc_state(self.h, t, ctypes.byref(s), ctypes.byref(st))
        return rc, s.value, st.value

    def stats(self):
        out = (i64 * 9)()
        self.L.sc_stats(self.h, out)
        keys = ("books", "queued", "running", "done", "failed", "cancelled", "launches",
                "options", "waits")
        return dict(zip(keys, out))


def test_layout_sizes(L):
    assert L.sc_book_bytes(1) == 6 * 256
    assert L.sc_book_bytes(1000) == 5 * 4096 + 1024
    assert L.sc_output_bytes(1000, PRICE) == 4096
    assert L.sc_output_bytes(1000, GREEKS) == 5 * 4096


def test_arena_against_a_model(L):
    """Random alloc/free against an explicit interval model: no overlap, every
    block inside the arena, used bytes exact, full coalescing when empty."""
    cap = 1 << 20
    a = L.sa_create(cap)
    rng = np.random.default_rng(5)
    live = {}
    for step in range(4000):
        if live and rng.random() < 0.45:
            off = int(rng.choice(list(live)))
            assert L.sa_free(a, off) == 1
            del live[off]
        else:
            size = int(rng.integers(1, 40000))
            o = i64()
            ok = L.sa_alloc(a, size, ctypes.byref(o))
            need = (size + 255) // 256 * 256
            if ok:
                off = o.value
                assert off % 256 == 0 and off + need <= cap
                for lo, sz in live.items():
                    assert off + need <= lo or lo + sz <= off, "overlap"
                live[off] = need
            else:
                assert L.sa_largest_free(a) < need
        assert L.sa_used(a) == sum(live.values())
        if step % 25 == 0:                              # largest free block == largest gap
            gaps, pos = [], 0
            for lo in sorted(live):
                gaps.append(lo - pos)
                pos = lo + live[lo]
            gaps.append(cap - pos)
            assert L.sa_largest_free(a) == max(gaps)
    assert L.sa_free(a, 12345) == 0                     # not a live block
    for off in list(live):
        L.sa_free(a, off)
    assert L.sa_used(a) == 0 and L.sa_largest_free(a) == cap
    L.sa_destroy(a)


def test_register_budget_and_release(L):
    c = Core(L, 1 << 20)
    rc, b1, _ = c.book(10_000)           # 5*40192 + 10240 bytes
    assert rc == OK
    assert c.book(100_000)[0] == BUDGET  # does not fit
    assert c.book(0)[0] == INVALID
    assert L.sc_release_book(c.h, 999) == NOT_FOUND
    rc, t = c.submit(b1)
    assert rc == OK
    assert L.sc_release_book(c.h, b1) == BUSY       # referenced by a queued request
    assert c.submit(b1, kind=7)[0] == INVALID
    assert c.submit(424242)[0] == NOT_FOUND
    tk, _ = c.admit()
    c.complete(tk)
    assert L.sc_release_book(c.h, b1) == OK


def test_fifo_admission_waits_for_memory(L):
    # arena: one book of 1000 options (21.5 KB) + room for ~3 price outputs
    bb = L.sc_book_bytes(1000)
    ob = L.sc_output_bytes(1000, PRICE)
    c = Core(L, bb + 3 * ob)
    _, b, _ = c.book(1000)
    tickets = [c.submit(b)[1] for _ in range(5)]
    first, offs = c.admit()
    assert first == tickets[:3] and len(set(offs)) == 3
    assert c.admit()[0] == []                        # no memory: the 4th waits
    assert c.stats()["waits"] >= 1
    assert c.state(tickets[3])[1] == QUEUED
    c.complete(first)
    assert c.admit()[0] == []                        # done but results not released
    assert L.sc_release_result(c.h, tickets[0]) == OK
    nxt, _ = c.admit()
    assert nxt == [tickets[3]]                       # FIFO order kept
    assert L.sc_release_result(c.h, tickets[1]) == OK
    assert c.admit()[0] == [tickets[4]]


def test_batch_limits(L):
    c = Core(L, 64 << 20, max_req=4, max_opts=2500)
    _, b, _ = c.book(1000)
    ts = [c.submit(b)[1] for _ in range(10)]
    got, _ = c.admit()
    assert got == ts[:2]                              # 2 x 1000 <= 2500 < 3 x 1000
    c2 = Core(L, 64 << 20, max_req=3)
    _, b2, _ = c2.book(10)
    ts2 = [c2.submit(b2)[1] for _ in range(7)]
    assert c2.admit()[0] == ts2[:3] and c2.admit()[0] == ts2[3:6]


def test_cancel_and_states(L):
    c = Core(L, 1 << 20)
    _, b, _ = c.book(100)
    _, t1 = c.submit(b)
    _, t2 = c.submit(b, GREEKS)
    assert L.sc_cancel(c.h, t2, 0) == OK
    assert c.state(t2)[1:] == (CANCELLED_ST, CANCELLED)
    assert L.sc_cancel(c.h, t2, 0) == BUSY           # no longer queued
    got, _ = c.admit()
    assert got == [t1]
    assert L.sc_cancel(c.h, t1, 0) == BUSY           # running requests finish
    assert L.sc_release_result(c.h, t1) == BUSY
    c.complete([t1])
    assert c.state(t1)[1] == DONE
    assert L.sc_release_result(c.h, t1) == OK
    assert c.state(t1)[0] == NOT_FOUND
    assert L.sc_release_result(c.h, t2) == OK


def test_failed_batch_frees_outputs(L):
    c = Core(L, 1 << 20)
    _, b, _ = c.book(1000)
    used0 = L.sc_arena_used(c.h)
    ts = [c.submit(b, GREEKS)[1] for _ in range(3)]
    got, _ = c.admit()
    assert L.sc_arena_used(c.h) == used0 + 3 * L.sc_output_bytes(1000, GREEKS)
    c.complete(got, ok=False)                        # injected launch failure
    for t in ts:
        assert c.state(t)[1:] == (FAILED_ST, FAILED)
    assert L.sc_arena_used(c.h) == used0
    assert L.sc_release_book(c.h, b) == OK           # no dangling references
    s = c.stats()
    assert s["failed"] == 3 and s["running"] == 0


def test_shutdown_fails_queued(L):
    c = Core(L, 1 << 20)
    _, b, _ = c.book(10)
    ts = [c.submit(b)[1] for _ in range(4)]
    L.sc_fail_queued(c.h, SHUTDOWN, 99)
    for t in ts:
        assert c.state(t)[1:] == (FAILED_ST, SHUTDOWN)
    assert L.sc_queued(c.h) == 0 and L.sc_release_book(c.h, b) == OK


def test_never_fitting_request_is_refused(L):
    # outputs that cannot fit next to their own book are refused at submit
    bb = L.sc_book_bytes(300_000)
    c = Core(L, bb + L.sc_output_bytes(300_000, PRICE))
    rc, b, _ = c.book(300_000)
    assert rc == OK
    assert c.submit(b, PRICE)[0] == OK
    assert c.submit(b, GREEKS)[0] == BUDGET


def reap(c, cap=64, users=False):
    t, s, st, u = (i64 * cap)(), (ctypes.c_int * cap)(), (ctypes.c_int * cap)(), (i64 * cap)()
    k = c.L.sc_reap(c.h, t, s, st, u, cap)
    if users:
        return [(t[i], u[i]) for i in range(k)]
    return [(t[i], s[i], st[i]) for i in range(k)]


def test_completion_queue_order_and_states(L):
    """Every finished request (done, failed, cancelled, shut down) is reported
    once, in the order it finished; released tickets are skipped."""
    c = Core(L, 1 << 22)
    L.sc_enable_completions(c.h, 1)
    _, b, _ = c.book(100)
    t = [c.submit(b)[1] for _ in range(6)]
    assert L.sc_cancel(c.h, t[1], 0) == OK
    got, _ = c.admit(cap=2)
    assert got == [t[0], t[2], t[3], t[4], t[5]][:len(got)]
    c.complete([t[2], t[0]])                           # completion order, not ticket order
    c.complete([x for x in got if x not in (t[0], t[2])], ok=False)
    assert reap(c, 1) == [(t[1], CANCELLED_ST, CANCELLED)]
    rest = reap(c)
    assert rest[:2] == [(t[2], DONE, OK), (t[0], DONE, OK)]
    assert all(s == FAILED_ST for _, s, _ in rest[2:])
    assert reap(c) == []
    # a ticket released before it is reaped is skipped
    _, t7 = c.submit(b)
    c.admit()
    c.complete([t7])
    assert L.sc_release_result(c.h, t7) == OK
    assert reap(c) == []
    _, t8 = c.submit(b)
    L.sc_fail_queued(c.h, SHUTDOWN, 5)
    assert reap(c) == [(t8, FAILED_ST, SHUTDOWN)]


def test_completion_queue_off_by_default(L):
    c = Core(L, 1 << 20)
    _, b, _ = c.book(10)
    _, t = c.submit(b)
    c.admit()
    c.complete([t])
    assert reap(c) == []


def test_submit_many_is_all_or_nothing(L):
    c = Core(L, 1 << 22)
    _, b1, _ = c.book(100)
    _, b2, _ = c.book(50)
    books = (i64 * 4)(b1, b2, b1, b2)
    kinds = (ctypes.c_int * 4)(PRICE, GREEKS, GREEKS, PRICE)
    tk = (i64 * 4)()
    L.sc_enable_completions(c.h, 1)
    users = (i64 * 4)(11, 22, 33, 44)
    assert L.sc_submit_many(c.h, books, kinds, users, 4, 1, tk) == OK
    assert len(set(tk)) == 4 and L.sc_queued(c.h) == 4
    got, _ = c.admit()
    assert got == list(tk)                             # queued in call order
    c.complete(got)
    assert reap(c, users=True) == list(zip(tk, [11, 22, 33, 44]))   # tags come back
    bad_book = (i64 * 3)(b1, 999, b2)
    assert L.sc_submit_many(c.h, bad_book, kinds, None, 3, 2, tk) == NOT_FOUND
    bad_kind = (ctypes.c_int * 3)(PRICE, 7, PRICE)
    assert L.sc_submit_many(c.h, books, bad_kind, None, 3, 2, tk) == INVALID
    assert L.sc_queued(c.h) == 0                       # nothing of a failed call queued
    assert L.sc_submit_many(c.h, books, kinds, None, 0, 2, tk) == OK


def test_admit_option_cap_and_queued_options(L):
    """A per-call option cap splits a long queue; the first request is
    admitted even when it alone exceeds the cap.  queued_options tracks the
    queue through submit, admit, cancel and shutdown."""
    c = Core(L, 1 << 24)
    _, b, _ = c.book(1000)
    _, big, _ = c.book(5000)
    ts = [c.submit(b)[1] for _ in range(6)]
    assert L.sc_queued_options(c.h) == 6000
    tk = (i64 * 16)()
    assert L.sc_admit_cap(c.h, 1, 3000, tk, 16) == 3 and list(tk[:3]) == ts[:3]
    assert L.sc_queued_options(c.h) == 3000
    _, tb = c.submit(big)
    assert L.sc_cancel(c.h, ts[4], 0) == OK
    assert L.sc_queued_options(c.h) == 7000
    assert L.sc_admit_cap(c.h, 2, 1500, tk, 16) == 1 and tk[0] == ts[3]
    assert L.sc_admit_cap(c.h, 3, 1500, tk, 16) == 1 and tk[0] == ts[5]
    assert L.sc_admit_cap(c.h, 4, 1500, tk, 16) == 1 and tk[0] == tb   # over the cap alone
    assert L.sc_queued_options(c.h) == 0
    c.submit(b)
    L.sc_fail_queued(c.h, SHUTDOWN, 9)
    assert L.sc_queued_options(c.h) == 0


def test_arena_size_classes_and_reuse(L):
    """Sizes across the allocator's classes (256 B to 64 MiB), including
    sizes just above a class edge, which the O(1) search skips: a hole
    between live blocks is found when it is the only block that fits (the
    exact search of the request's own class); neighbours coalesce in either
    order; an exactly filled arena refuses even one byte."""
    cap = 1 << 30
    a = L.sa_create(cap)
    o = [i64() for _ in range(5)]
    sizes = [256, 4096, 15 * 256, 16 * 256, 17 * 256, 1 << 20, (1 << 20) + 256,
             (1 << 20) + 255 * 256, 64 << 20]
    for sz in sizes:
        assert L.sa_alloc(a, sz, ctypes.byref(o[0])) == 1
        assert L.sa_alloc(a, sz, ctypes.byref(o[1])) == 1
        assert L.sa_alloc(a, sz, ctypes.byref(o[2])) == 1
        rest = L.sa_largest_free(a)
        assert rest == cap - 3 * ((sz + 255) // 256 * 256)
        assert L.sa_alloc(a, rest, ctypes.byref(o[4])) == 1    # fill the arena
        assert L.sa_largest_free(a) == 0
        assert L.sa_free(a, o[1].value) == 1                   # hole between live blocks
        assert L.sa_alloc(a, sz + 256, ctypes.byref(o[3])) == 0  # one unit too big
        assert L.sa_alloc(a, sz, ctypes.byref(o[3])) == 1
        assert o[3].value == o[1].value                        # the only fit is used
        for x in (o[0], o[4], o[2], o[3]):
            assert L.sa_free(a, x.value) == 1
        assert L.sa_used(a) == 0 and L.sa_largest_free(a) == cap
    # fill exactly, then free in interleaved order
    offs = []
    for _ in range(16):
        assert L.sa_alloc(a, cap // 16, ctypes.byref(o[0])) == 1
        offs.append(o[0].value)
    assert L.sa_largest_free(a) == 0 and L.sa_alloc(a, 1, ctypes.byref(o[1])) == 0
    for off in offs[::2] + offs[1::2]:
        assert L.sa_free(a, off) == 1
    assert L.sa_largest_free(a) == cap
    L.sa_destroy(a)


def test_small_requests_have_their_own_lane(L):
    """With small_options = 1000: requests of <= 1000 options queue in lane 0
    and are admitted only into lane-0 batches, FIFO within the lane; a bulk
    request waiting for memory does not hold them back; cancel and shutdown
    reach both lanes."""
    book_n = {}
    # arena: books (small 1000, big 100000) plus outputs of one big Greeks request
    cap = 256 * 1024 * 1024
    h = L.sc_create_lanes(cap, 1000)
    try:
        ids = {}
        for n in (10, 1000, 1001, 100_000):
            b, off = i64(), i64()
            assert L.sc_register(h, n, ctypes.byref(b), ctypes.byref(off)) == OK
            ids[n] = b.value
            book_n[b.value] = n
        t = i64()
        order = []
        for n in (100_000, 10, 1001, 1000, 10, 100_000):
            assert L.sc_submit(h, ids[n], PRICE, 0, ctypes.byref(t)) == OK
            order.append((n, t.value))
        assert L.sc_queued_lane(h, 0) == 3 and L.sc_queued_lane(h, 1) == 3
        tk = (i64 * 16)()
        k = L.sc_admit_lane(h, 1, 0, tk, 16)
        assert list(tk[:k]) == [tv for n, tv in order if n <= 1000]       # FIFO in lane 0
        k = L.sc_admit_lane(h, 2, 1, tk, 16)
        assert list(tk[:k]) == [tv for n, tv in order if n > 1000]        # the bulk lane's own
        assert L.sc_queued_lane(h, 0) == 0 and L.sc_queued_lane(h, 1) == 0
        # cancel in the small lane, shutdown with both lanes queued
        assert L.sc_submit(h, ids[10], PRICE, 0, ctypes.byref(t)) == OK
        small = t.value
        assert L.sc_cancel(h, small, 0) == OK and L.sc_queued_lane(h, 0) == 0
        for n in (10, 100_000):
            assert L.sc_submit(h, ids[n], GREEKS, 0, ctypes.byref(t)) == OK
        L.sc_fail_queued(h, SHUTDOWN, 5)
        assert L.sc_queued_lane(h, 0) == 0 and L.sc_queued_lane(h, 1) == 0
    finally:
        L.sc_destroy(h)


def test_bulk_head_waiting_for_memory_does_not_block_the_small_lane(L):
    # the arena holds the books and one big Greeks output; a second big request waits
    small_n, big_n = 100, 200_000
    cap = 5 * (4 * big_n + 256) + 20 * 1024 + 5 * 4 * big_n * 2 - 4096
    h = L.sc_create_lanes(cap, 1000)
    try:
        b, off, t = i64(), i64(), i64()
        assert L.sc_register(h, big_n, ctypes.byref(b), ctypes.byref(off)) == OK
        big = b.value
        assert L.sc_register(h, small_n, ctypes.byref(b), ctypes.byref(off)) == OK
        sm = b.value
        tk = (i64 * 16)()
        for _ in range(2):
            assert L.sc_submit(h, big, GREEKS, 0, ctypes.byref(t)) == OK
        assert L.sc_admit_lane(h, 1, 1, tk, 16) == 1      # the second big one does not fit
        assert L.sc_submit(h, sm, PRICE, 0, ctypes.byref(t)) == OK
        assert L.sc_admit_lane(h, 2, 1, tk, 16) == 0      # bulk head still waits
        assert L.sc_admit_lane(h, 3, 0, tk, 16) == 1 and tk[0] == t.value   # small lane goes on
    finally:
        L.sc_destroy(h)
