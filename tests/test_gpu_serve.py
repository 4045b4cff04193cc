"""The serving loop on the MI355X, through the C ABI.

* every result of the fused, graph-replayed launch is bit-identical to the
  standalone map (bs_price_f32 / bs_price_greeks_f32) -- alone and in a mix;
* admission by memory: with a small arena, queued requests wait and are
  admitted as results are released; all complete;
* cancellation, budget refusal, busy books, shutdown with pending work;
* bulk submission and the completion queue: every request reported once,
  results identical to the standalone map, memory returned by bulk release.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu
i64 = ctypes.c_int64


class Ctx:
    def __init__(self, capi, budget, use_graph=1, max_req=0, grid=0, cq=0, small=0):
        self.capi = capi
        self.L = capi.load()
        o = capi.ServeOpts(max_req, 0, use_graph, grid, cq, small)
        h = ctypes.c_void_p()
        capi.call("bs_ctx_create", 0, budget, ctypes.byref(o), ctypes.byref(h))
        self.h = h.value

    def register(self, book):
        b = i64()
        arrs = [np.ascontiguousarray(a) for a in book.arrays()]
        self.capi.call("bs_book_register", self.h, *[a.ctypes.data for a in arrs], book.n,
                       ctypes.byref(b))
        return b.value

    def submit(self, book, kind):
        t = i64()
        self.capi.call("bs_submit", self.h, book, kind, ctypes.byref(t))
        return t.value

    def wait(self, t, timeout_us=30_000_000):
        return self.L.bs_wait(self.h, t, timeout_us)

    def result(self, t, n, kind):
        outs = [np.empty(n, np.float32) for _ in range(5 if kind == 2 else 1)]
        ptrs = [o.ctypes.data for o in outs] + [None] * (5 - len(outs))
        self.capi.call("bs_result_copy", self.h, t, *ptrs)
        return outs

    def stats(self):
        s = self.capi.ServeStats()
        self.capi.call("bs_ctx_stats", self.h, ctypes.byref(s))
        return s

    def destroy(self):
        if self.h:
            self.capi.call("bs_ctx_destroy", self.h)
            self.h = None


def standalone(capi, book, kind):
    db = capi.DeviceBook(book)
    if kind == 1:
        out = [capi.price(db).download(np.float32, book.n)]
    else:
        out = [o.download(np.float32, book.n) for o in capi.greeks(db)]
    db.free()
    return out


def same_bits(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


@pytest.mark.parametrize("use_graph,small", [(1, 0), (0, 0), (0, -1), (1, 1024)])
def test_alone_and_in_the_mix(gpu, use_graph, small):
    """Each result bit-identical to the batch map, alone and in a 200-request
    mix; small = 0 puts books of <= 64K options in the small lane, 1024 a
    lower threshold, -1 one lane for all."""
    rng = np.random.default_rng(7)
    sizes = [1, 3, 4, 5, 1023, 1024, 1025, 4096, 70_001, 262_144]
    bks = [books.make_book(n, books.SEED_BASE + 40 + i) for i, n in enumerate(sizes)]
    ctx = Ctx(gpu, 512 << 20, use_graph=use_graph, small=small)
    try:
        ids = [ctx.register(b) for b in bks]
        ref = {(i, k): standalone(gpu, bks[i], k) for i in range(len(bks)) for k in (1, 2)}
        # alone: one request at a time
        for i in range(len(bks)):
            for k in (1, 2):
                t = ctx.submit(ids[i], k)
                assert ctx.wait(t) == 0
                assert same_bits(ctx.result(t, bks[i].n, k), ref[(i, k)]), (i, k)
                gpu.call("bs_result_release", ctx.h, t)
        # in the mix: 200 requests submitted back to back, random books and kinds
        reqs = [(int(rng.integers(len(bks))), int(rng.integers(1, 3))) for _ in range(200)]
        ts = [ctx.submit(ids[i], k) for i, k in reqs]
        for t, (i, k) in zip(ts, reqs):
            assert ctx.wait(t) == 0
            assert same_bits(ctx.result(t, bks[i].n, k), ref[(i, k)]), (i, k)
            gpu.call("bs_result_release", ctx.h, t)
        st = ctx.stats()
        print(f"\nuse_graph={use_graph} small={small}: launches {st.launches} for "
              f"{st.requests_done} requests")
        assert st.requests_done == 2 * len(bks) + 200 and st.requests_failed == 0
        assert st.launches < st.requests_done        # requests were fused
    finally:
        ctx.destroy()


@pytest.mark.parametrize("threads,use_graph,grid", [(64, 0, 0), (64, 1, 0), (128, 0, 0),
                                                    (128, 1, 5), (64, 0, 5)])
def test_block_threads_alone_and_in_the_mix(gpu, threads, use_graph, grid):
    """BS_CFG_BLOCK_THREADS at context creation sets the fused kernels' block
    (a tile is threads x 2 options): results stay bit-identical to the batch
    maps priced with the default launch shape, alone and in a 120-request mix,
    in both launch forms and with a capped grid (blocks loop over tiles)."""
    rng = np.random.default_rng(threads + grid)
    sizes = [1, 2, 127, 128, 129, 255, 256, 257, 4097, 70_001, 262_144]
    bks = [books.make_book(n, books.SEED_BASE + 90 + i) for i, n in enumerate(sizes)]
    ref = {(i, k): standalone(gpu, bks[i], k) for i in range(len(bks)) for k in (1, 2)}
    gpu.call("bs_config_set", gpu.CFG_BLOCK_THREADS, threads)
    try:
        ctx = Ctx(gpu, 512 << 20, use_graph=use_graph, grid=grid)
    finally:
        gpu.call("bs_config_set", gpu.CFG_BLOCK_THREADS, 0)
    try:
        ids = [ctx.register(b) for b in bks]
        reqs = [(i, k) for i in range(len(bks)) for k in (1, 2)]
        for i, k in reqs:                                   # alone
            t = ctx.submit(ids[i], k)
            assert ctx.wait(t) == 0
            assert same_bits(ctx.result(t, bks[i].n, k), ref[(i, k)]), (threads, i, k)
            gpu.call("bs_result_release", ctx.h, t)
        reqs = [(int(rng.integers(len(bks))), int(rng.integers(1, 3))) for _ in range(120)]
        ts = [ctx.submit(ids[i], k) for i, k in reqs]       # in the mix
        for t, (i, k) in zip(ts, reqs):
            assert ctx.wait(t) == 0
            assert same_bits(ctx.result(t, bks[i].n, k), ref[(i, k)]), (threads, i, k)
            gpu.call("bs_result_release", ctx.h, t)
        st = ctx.stats()
        assert st.requests_failed == 0 and st.launches < st.requests_done
    finally:
        ctx.destroy()


@pytest.mark.parametrize("use_graph", [1, 0])
def test_capped_grid_loops_over_tiles(gpu, use_graph):
    """grid_blocks = 7: each section's blocks loop over the batch's tiles;
    results stay bit-identical to the batch maps, and the graph form's section
    grids follow each batch's size (1 tile, then hundreds)."""
    sizes = [1, 513, 70_001, 262_144]
    bks = [books.make_book(n, books.SEED_BASE + 70 + i) for i, n in enumerate(sizes)]
    ctx = Ctx(gpu, 256 << 20, use_graph=use_graph, grid=7)
    try:
        ids = [ctx.register(b) for b in bks]
        ref = {(i, k): standalone(gpu, bks[i], k) for i in range(len(bks)) for k in (1, 2)}
        reqs = [(i, k) for k in (1, 2) for i in range(len(bks))] * 3
        for i, k in reqs[:8]:      # alone
            t = ctx.submit(ids[i], k)
            assert ctx.wait(t) == 0
            assert same_bits(ctx.result(t, bks[i].n, k), ref[(i, k)]), (i, k)
            gpu.call("bs_result_release", ctx.h, t)
        ts = [ctx.submit(ids[i], k) for i, k in reqs]   # in the mix
        for t, (i, k) in zip(ts, reqs):
            assert ctx.wait(t) == 0
            assert same_bits(ctx.result(t, bks[i].n, k), ref[(i, k)]), (i, k)
            gpu.call("bs_result_release", ctx.h, t)
        assert ctx.stats().requests_failed == 0
    finally:
        ctx.destroy()


def test_admission_by_memory(gpu):
    n = 100_000
    book = books.make_book(n, books.SEED_BASE + 50)
    ref = standalone(gpu, book, 2)
    from p3arsec_amd.capi import load  # noqa: F401
    # arena: the book + outputs of ~3 Greeks requests
    budget = 5 * 400_128 + 100_096 + 3 * 5 * 400_128 + 4096
    ctx = Ctx(gpu, budget)
    try:
        b = ctx.register(book)
        ts = [ctx.submit(b, 2) for _ in range(12)]
        done = []

        def consumer():
            for t in ts:
                assert ctx.wait(t) == 0
                done.append(ctx.result(t, n, 2))
                gpu.call("bs_result_release", ctx.h, t)

        th = threading.Thread(target=consumer)
        th.start()
        th.join(60)
        assert not th.is_alive()
        assert len(done) == 12 and all(same_bits(d, ref) for d in done)
        st = ctx.stats()
        assert st.admission_waits > 0 and st.arena_peak <= st.arena_bytes
    finally:
        ctx.destroy()


def test_cancel_budget_busy_and_errors(gpu):
    n = 50_000
    book = books.make_book(n, books.SEED_BASE + 51)
    budget = 5 * 200_192 + 50_176 + 200_192       # book + one price output
    ctx = Ctx(gpu, budget)
    L = ctx.L
    try:
        b = ctx.register(book)
        t1 = ctx.submit(b, 1)
        assert ctx.wait(t1) == 0                      # holds the only output slot
        t2 = ctx.submit(b, 1)                         # must wait for memory
        time.sleep(0.05)
        st = ctypes.c_int()
        gpu.call("bs_poll", ctx.h, t2, ctypes.byref(st))
        assert st.value == gpu.TICKET_QUEUED
        assert L.bs_cancel(ctx.h, t2) == 0
        assert ctx.wait(t2) == -6                     # BS_ERR_CANCELLED
        assert L.bs_submit(ctx.h, b, 2, ctypes.byref(i64())) == -4   # Greeks never fit
        assert L.bs_book_release(ctx.h, b) == 0        # t1 done, t2 cancelled: no refs
        big = books.make_book(200_000, books.SEED_BASE + 52)
        arrs = [np.ascontiguousarray(a) for a in big.arrays()]
        assert L.bs_book_register(ctx.h, *[a.ctypes.data for a in arrs], big.n,
                                  ctypes.byref(i64())) == -4          # BS_ERR_BUDGET
        assert L.bs_submit(ctx.h, 9999, 1, ctypes.byref(i64())) == -5  # unknown book
        assert L.bs_wait(ctx.h, 123456, 0) == -5
        gpu.call("bs_result_release", ctx.h, t1)
    finally:
        ctx.destroy()


def test_busy_book_and_shutdown_with_pending(gpu):
    book = books.make_book(1000, books.SEED_BASE + 53)
    ctx = Ctx(gpu, 64 << 20)
    b = ctx.register(book)
    ts = [ctx.submit(b, 1) for _ in range(50)]
    rc = ctx.L.bs_book_release(ctx.h, b)
    assert rc in (0, -8)                              # BS_ERR_BUSY while referenced
    ctx.destroy()                                     # drains or fails pending work; no hang
    assert ts


def test_resident_book_repriced(gpu):
    book = books.make_book(300_000, books.SEED_BASE + 54)
    ref = standalone(gpu, book, 1)
    ctx = Ctx(gpu, 256 << 20)
    try:
        b = ctx.register(book)
        for _ in range(20):
            t = ctx.submit(b, 1)
            assert ctx.wait(t) == 0
            assert same_bits(ctx.result(t, book.n, 1), ref)
            gpu.call("bs_result_release", ctx.h, t)
        s0, s1 = i64(), i64()
        gpu.call("bs_ticket_times", ctx.h, ctx.submit(b, 1), ctypes.byref(s0), ctypes.byref(s1))
        assert s0.value > 0
    finally:
        ctx.destroy()


def test_submit_many_and_completion_queue(gpu):
    rng = np.random.default_rng(11)
    sizes = [1, 7, 1024, 5000, 65_537, 200_000]
    bks = [books.make_book(n, books.SEED_BASE + 70 + i) for i, n in enumerate(sizes)]
    ref = {(i, k): standalone(gpu, bks[i], k) for i in range(len(bks)) for k in (1, 2)}
    # the arena holds the books and about 6 Greeks outputs of the largest book,
    # so bulk submissions wait for memory that only bulk releases return
    book_bytes = sum(5 * (((n * 4 + 255) // 256) * 256) + ((n + 255) // 256) * 256
                     for n in sizes)
    budget = book_bytes + 6 * 5 * ((200_000 * 4 + 255) // 256) * 256
    ctx = Ctx(gpu, budget, cq=1)
    L = ctx.L
    try:
        ids = [ctx.register(b) for b in bks]
        reqs = [(int(rng.integers(len(bks))), int(rng.integers(1, 3))) for _ in range(300)]
        kinds = {}
        for lo in range(0, len(reqs), 100):
            chunk = reqs[lo:lo + 100]
            bk = (i64 * len(chunk))(*[ids[i] for i, _ in chunk])
            kd = (ctypes.c_int * len(chunk))(*[k for _, k in chunk])
            tk = (i64 * len(chunk))()
            us = (i64 * len(chunk))(*range(lo, lo + len(chunk)))
            gpu.call("bs_submit_many", ctx.h, bk, kd, us, len(chunk), tk)
            kinds.update({t: (r, u) for t, r, u in zip(tk, chunk, us)})
        cq = (gpu.Completion * 64)()
        cnt = i64()
        seen = []
        deadline = time.time() + 60
        while len(seen) < len(reqs) and time.time() < deadline:
            gpu.call("bs_reap", ctx.h, cq, 64, 100_000, ctypes.byref(cnt))
            got = [cq[j] for j in range(cnt.value)]
            for c in got:
                assert c.state == gpu.TICKET_DONE and c.status == 0
                assert c.done_ns >= c.submit_ns > 0
                (i, k), u = kinds[c.ticket]
                assert c.user == u and c.n == bks[i].n
                assert same_bits(ctx.result(c.ticket, bks[i].n, k), ref[(i, k)]), (i, k)
            rel = (i64 * max(1, len(got)))(*[c.ticket for c in got])
            gpu.call("bs_result_release_many", ctx.h, rel, len(got))
            seen += [c.ticket for c in got]
        assert sorted(seen) == sorted(kinds)          # each request reported once
        st = ctx.stats()
        assert st.requests_done == len(reqs) and st.admission_waits > 0
        assert st.arena_used == book_bytes            # every output returned
        # errors: all-or-nothing submit, double release, reap with max 0
        bk = (i64 * 2)(ids[0], 424242)
        kd = (ctypes.c_int * 2)(1, 1)
        assert L.bs_submit_many(ctx.h, bk, kd, None, 2, (i64 * 2)()) == -5
        gpu.call("bs_reap", ctx.h, cq, 64, 0, ctypes.byref(cnt))
        assert cnt.value == 0                         # the failed call queued nothing
        assert L.bs_result_release_many(ctx.h, (i64 * 1)(seen[0]), 1) == -5
    finally:
        ctx.destroy()


def test_reap_needs_completion_queue(gpu):
    ctx = Ctx(gpu, 16 << 20)
    try:
        cnt = i64()
        assert ctx.L.bs_reap(ctx.h, (gpu.Completion * 4)(), 4, 0, ctypes.byref(cnt)) == -1
    finally:
        ctx.destroy()
