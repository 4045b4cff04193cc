"""The entry points inside a caller's own hipGraph (MI355X): a caller captures
bs_price_f32 and bs_price_greeks_f32 on its stream, instantiates the graph
and replays it.  Every replay must equal the eager result for the inputs the
replay saw: the graph is replayed six times with the spot array rewritten and
eager passes (on another output) between replays.  With the library's
pointer checks on (the default) and off.  The HIP runtime is called directly
(ctypes), as the caller would."""
import ctypes

import numpy as np
import pytest

from p3arsec_amd import books

pytestmark = pytest.mark.gpu
CAPTURE_THREAD_LOCAL = 1


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    h.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
    h.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
    h.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]
    h.hipGraphLaunch.argtypes = [vp, vp]
    h.hipGraphExecDestroy.argtypes = [vp]
    h.hipGraphDestroy.argtypes = [vp]
    for f in ("hipStreamBeginCapture", "hipStreamEndCapture", "hipGraphInstantiate",
              "hipGraphLaunch", "hipGraphExecDestroy", "hipGraphDestroy"):
        getattr(h, f).restype = ctypes.c_int
    return h


@pytest.mark.parametrize("check_pointers", [1, 0])
def test_captured_in_a_callers_graph(gpu, check_pointers):
    H = hip()
    L = gpu.load()
    n = 100_003
    base = books.make_book(n, books.SEED_BASE + 140)
    db = gpu.DeviceBook(base)
    out = gpu.alloc_out(n)
    gouts = [gpu.alloc_out(n) for _ in range(5)]
    st = gpu.Stream()
    old = gpu.get_config(gpu.CFG_CHECK_POINTERS)
    gpu.set_config(gpu.CFG_CHECK_POINTERS, check_pointers)
    graph, gexec = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        assert H.hipStreamBeginCapture(st.handle, CAPTURE_THREAD_LOCAL) == 0
        rc1 = L.bs_price_f32(*db.ptrs(), out.ptr, n, st.handle)
        rc2 = L.bs_price_greeks_f32(*db.ptrs(), *[o.ptr for o in gouts], n, st.handle)
        assert H.hipStreamEndCapture(st.handle, ctypes.byref(graph)) == 0
        assert rc1 == 0 and rc2 == 0, L.bs_last_error()
        assert H.hipGraphInstantiate(ctypes.byref(gexec), graph, None, None, 0) == 0
        rng = np.random.default_rng(9)
        for rep in range(6):
            s = (base.s * np.float32(rng.uniform(0.8, 1.2))).astype(np.float32)
            gpu.call("bs_memcpy_h2d", db.ptrs()[0], s.ctypes.data, n * 4, None)
            gpu.call("bs_stream_sync", None)
            want = gpu.price(db).download(np.float32, n)            # eager, another output
            want_g = [o.download(np.float32, n) for o in gpu.greeks(db)]
            assert H.hipGraphLaunch(gexec, st.handle) == 0
            gpu.call("bs_stream_sync", st.handle)
            assert np.array_equal(out.download(np.float32, n).view(np.uint32),
                                  want.view(np.uint32)), rep
            for o, w in zip(gouts, want_g):
                assert np.array_equal(o.download(np.float32, n).view(np.uint32),
                                      w.view(np.uint32)), rep
    finally:
        gpu.set_config(gpu.CFG_CHECK_POINTERS, old)
        if gexec.value:
            H.hipGraphExecDestroy(gexec)
        if graph.value:
            H.hipGraphDestroy(graph)
        st.destroy()
        for o in [out] + gouts:
            o.free()
        db.free()
