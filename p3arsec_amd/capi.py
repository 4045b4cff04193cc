"""Python binding of libbs_mi355.so (ctypes over the C ABI in include/bs_mi355.h).

This is a thin front for tests, benchmarks and Python callers: every compute
call goes to the HIP kernels of the shared library.  There is no CPU path --
if the library or a GPU is missing, calls raise.

Device memory is handled through the library's own plumbing entry points
(bs_malloc, bs_memcpy_*, streams and events), so the binding needs neither
PyTorch nor any other GPU runtime.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libbs_mi355.so")
HEADER = os.path.join(os.path.dirname(_PKG), "include", "bs_mi355.h")

BS_OK = 0
STATUS = {
    0: "BS_OK", -1: "BS_ERR_INVALID", -2: "BS_ERR_HIP", -3: "BS_ERR_NOMEM",
    -4: "BS_ERR_BUDGET", -5: "BS_ERR_NOT_FOUND", -6: "BS_ERR_CANCELLED",
    -7: "BS_ERR_TIMEOUT", -8: "BS_ERR_BUSY", -9: "BS_ERR_NODEVICE",
    -10: "BS_ERR_FAILED", -11: "BS_ERR_SHUTDOWN",
}
CFG_CHECK_POINTERS, CFG_MAX_BLOCKS, CFG_BLOCK_THREADS, CFG_VECTOR_WIDTH, CFG_NONTEMPORAL = 1, 2, 3, 4, 5
CFG_PIPELINE = 6
CFG_LDS_PER_WAVE = 7
REQ_PRICE, REQ_GREEKS = 1, 2
TICKET_QUEUED, TICKET_RUNNING, TICKET_DONE, TICKET_FAILED, TICKET_CANCELLED = range(5)


class BSError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_sz = ctypes.c_size_t
_f32 = ctypes.c_float


class ServeOpts(ctypes.Structure):
    _fields_ = [("max_batch_requests", _i64), ("max_batch_options", _i64),
                ("use_graph", ctypes.c_int32), ("grid_blocks", ctypes.c_int32),
                ("completion_queue", ctypes.c_int32), ("small_request_options", ctypes.c_int32)]


class Completion(ctypes.Structure):
    _fields_ = [("ticket", _i64), ("user", _i64), ("n", _i64), ("state", ctypes.c_int32),
                ("status", ctypes.c_int32), ("submit_ns", _i64), ("done_ns", _i64)]


class ServeStats(ctypes.Structure):
    _fields_ = [(n, _i64) for n in (
        "arena_bytes", "arena_used", "arena_peak", "books_resident", "requests_queued",
        "requests_running", "requests_done", "requests_failed", "requests_cancelled",
        "launches", "options_priced", "admission_waits", "queue_ns_total",
        "batch_ns_total", "dispatch_admit_ns", "dispatch_launch_ns", "dispatch_retire_ns",
        "gpu_idle_ns", "idle_launches")]


class Swaption(ctypes.Structure):
    """bs_swaption_t: one swaption's parameters (host pointers)."""
    _fields_ = [("strike", ctypes.c_double), ("compounding", ctypes.c_double),
                ("maturity", ctypes.c_double), ("tenor", ctypes.c_double),
                ("payment_interval", ctypes.c_double), ("years", ctypes.c_double),
                ("n_steps", ctypes.c_int32), ("n_factors", ctypes.c_int32),
                ("yield_", ctypes.POINTER(ctypes.c_double)),
                ("factors", ctypes.POINTER(ctypes.c_double))]


# name -> (argtypes, restype); restype int unless given
_SIGS = {
    "bs_status_string": ([_i32], ctypes.c_char_p),
    "bs_last_error": ([], ctypes.c_char_p),
    "bs_version": ([], _i32),
    "bs_config_set": ([_i32, _i64], _i32),
    "bs_config_get": ([_i32, ctypes.POINTER(_i64)], _i32),
    "bs_cndf_f32": ([_vp, _vp, _i64, _vp], _i32),
    "bs_cndf_f64": ([_vp, _vp, _i64, _vp], _i32),
    "bs_cndf_scalar_f32": ([_f32, ctypes.POINTER(_f32)], _i32),
    "bs_price_scalar_f32": ([_f32] * 5 + [_i32, _f32, ctypes.POINTER(_f32)], _i32),
    "bs_price_f32": ([_vp] * 7 + [_i64, _vp], _i32),
    "bs_price_i32type_f32": ([_vp] * 7 + [_i64, _vp], _i32),
    "bs_price_passes_f32": ([_vp] * 6 + [_i32, _vp, _i64, _i32, _vp], _i32),
    "bs_price_fast_f32": ([_vp] * 7 + [_i64, _vp], _i32),
    "bs_price_f64": ([_vp] * 7 + [_i64, _vp], _i32),
    "bs_price_greeks_f32": ([_vp] * 11 + [_i64, _vp], _i32),
    "bs_repack_aos_f32": ([_vp, _i64] + [_vp] * 6 + [_vp], _i32),
    "bs_price_host_f32": ([_vp] * 7 + [_i64, _i32, _i32], _i32),
    "bs_bind_globals": ([_i32, _i32] + [_vp] * 7, _i32),
    "bs_thread_compat": ([_vp], _i32),
    "bs_device_count": ([ctypes.POINTER(_i32)], _i32),
    "bs_set_device": ([_i32], _i32),
    "bs_malloc": ([ctypes.POINTER(_vp), _sz], _i32),
    "bs_free": ([_vp], _i32),
    "bs_memcpy_h2d": ([_vp, _vp, _sz, _vp], _i32),
    "bs_memcpy_d2h": ([_vp, _vp, _sz, _vp], _i32),
    "bs_memcpy_d2d": ([_vp, _vp, _sz, _vp], _i32),
    "bs_memset": ([_vp, _i32, _sz, _vp], _i32),
    "bs_stream_create": ([ctypes.POINTER(_vp)], _i32),
    "bs_stream_destroy": ([_vp], _i32),
    "bs_stream_sync": ([_vp], _i32),
    "bs_portfolio_f32": ([_vp] * 7 + [_i64, _i32, _vp, _vp], _i32),
    "bs_portfolio_f32_noalloc": ([_vp] * 7 + [_i64, _i32, _vp, _vp], _i32),
    "bs_stream_wait_event": ([_vp, _vp], _i32),
    "bs_device_sync": ([], _i32),
    "bs_event_create": ([ctypes.POINTER(_vp)], _i32),
    "bs_event_create_timing": ([ctypes.POINTER(_vp)], _i32),
    "bs_event_destroy": ([_vp], _i32),
    "bs_event_record": ([_vp, _vp], _i32),
    "bs_event_sync": ([_vp], _i32),
    "bs_event_elapsed_ms": ([_vp, _vp, ctypes.POINTER(_f32)], _i32),
    "bs_ctx_create": ([_i32, _sz, ctypes.POINTER(ServeOpts), ctypes.POINTER(_vp)], _i32),
    "bs_ctx_destroy": ([_vp], _i32),
    "bs_book_register": ([_vp] + [_vp] * 6 + [_i64, ctypes.POINTER(_i64)], _i32),
    "bs_book_release": ([_vp, _i64], _i32),
    "bs_submit": ([_vp, _i64, _i32, ctypes.POINTER(_i64)], _i32),
    "bs_submit_many": ([_vp, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(_i64), _i64, ctypes.POINTER(_i64)], _i32),
    "bs_reap": ([_vp, ctypes.POINTER(Completion), _i64, _i64, ctypes.POINTER(_i64)], _i32),
    "bs_result_release_many": ([_vp, ctypes.POINTER(_i64), _i64], _i32),
    "bs_poll": ([_vp, _i64, ctypes.POINTER(_i32)], _i32),
    "bs_wait": ([_vp, _i64, _i64], _i32),
    "bs_result_copy": ([_vp, _i64] + [_vp] * 5, _i32),
    "bs_result_device": ([_vp, _i64] + [ctypes.POINTER(_vp)] * 5, _i32),
    "bs_result_release": ([_vp, _i64], _i32),
    "bs_cancel": ([_vp, _i64], _i32),
    "bs_ticket_times": ([_vp, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)], _i32),
    "bs_ctx_stats": ([_vp, ctypes.POINTER(ServeStats)], _i32),
    "bs_swaptions_host_f64": ([ctypes.POINTER(Swaption), _i64, _i64, _i64, ctypes.c_int32,
                               _vp, _vp, _vp], _i32),
}

_lib = None


def header_functions(path: str = HEADER) -> list[str]:
    """Names of the functions include/bs_mi355.h declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bs_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load libbs_mi355.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `make -C p3arsec_amd/csrc`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (argt, rest) in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:      # test_capi_cpu checks that every declared symbol exists
            continue
        fn.argtypes = argt
        fn.restype = rest
    _lib = L
    return L


def check(rc: int, where: str) -> None:
    if rc != BS_OK:
        msg = load().bs_last_error()
        raise BSError(rc, where, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


# ---------------------------------------------------------------------------
def device_count() -> int:
    n = _i32(0)
    rc = load().bs_device_count(ctypes.byref(n))
    return n.value if rc == BS_OK else 0


def set_config(key: int, value: int) -> None:
    call("bs_config_set", key, value)


def get_config(key: int) -> int:
    v = _i64(0)
    call("bs_config_get", key, ctypes.byref(v))
    return v.value


class DeviceBuffer:
    """An owned device allocation (bs_malloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        call("bs_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            call("bs_memcpy_h2d", self.ptr, a.ctypes.data, a.nbytes, stream)
            if stream is None:
                call("bs_stream_sync", None)

    def download(self, dtype, count: int, stream=None, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset_bytes + out.nbytes <= self.nbytes
        if out.nbytes:
            call("bs_memcpy_d2h", out.ctypes.data, self.ptr + offset_bytes, out.nbytes, stream)
            call("bs_stream_sync", stream)
        return out

    def free(self) -> None:
        if self.ptr:
            call("bs_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.bs_free(self.ptr)
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = _vp()
        call("bs_stream_create", ctypes.byref(p))
        self.handle = p.value

    def sync(self) -> None:
        call("bs_stream_sync", self.handle)

    def wait(self, ev: "Event") -> None:
        call("bs_stream_wait_event", self.handle, ev.handle)

    def destroy(self) -> None:
        if self.handle:
            call("bs_stream_destroy", self.handle)
            self.handle = None


class Event:
    """A HIP event; timing=True makes a timing-only event (bs_event_create_timing)."""

    def __init__(self, timing: bool = False):
        p = _vp()
        call("bs_event_create_timing" if timing else "bs_event_create", ctypes.byref(p))
        self.handle = p.value

    def record(self, stream=None) -> None:
        call("bs_event_record", self.handle, stream)

    def sync(self) -> None:
        call("bs_event_sync", self.handle)

    def elapsed_ms(self, end: "Event") -> float:
        ms = _f32(0)
        call("bs_event_elapsed_ms", self.handle, end.handle, ctypes.byref(ms))
        return ms.value

    def destroy(self) -> None:
        if self.handle:
            call("bs_event_destroy", self.handle)
            self.handle = None


class DeviceBook:
    """A book resident in device memory: one buffer per SoA field."""

    def __init__(self, book, dtype=np.float32, type_dtype=np.uint8, pad: int = 0):
        self.n = book.n
        self.dtype = np.dtype(dtype)
        self.pad = pad  # element offset into each buffer (misalignment tests)
        self.bufs = []
        for a in (book.s, book.k, book.r, book.v, book.t):
            arr = np.concatenate([np.zeros(pad, self.dtype), a.astype(self.dtype)])
            self.bufs.append(DeviceBuffer.from_array(arr))
        ty = np.concatenate([np.zeros(pad, type_dtype), book.otype.astype(type_dtype)])
        self.type_buf = DeviceBuffer.from_array(ty)
        self.type_dtype = np.dtype(type_dtype)

    def ptrs(self):
        off = self.pad * self.dtype.itemsize
        return [b.ptr + off for b in self.bufs] + [self.type_buf.ptr +
                                                   self.pad * self.type_dtype.itemsize]

    def free(self):
        for b in self.bufs + [self.type_buf]:
            b.free()


def alloc_out(n: int, dtype=np.float32, pad: int = 0) -> DeviceBuffer:
    return DeviceBuffer((n + pad) * np.dtype(dtype).itemsize)


def price(dbook: DeviceBook, out: DeviceBuffer | None = None, mode: str = "ref",
          stream=None, sync: bool = True, out_pad: int = 0) -> DeviceBuffer:
    """Launch the price map over a device book; mode: ref | fast | i32 | f64."""
    dt = np.float64 if mode == "f64" else np.float32
    if out is None:
        out = alloc_out(dbook.n, dt, out_pad)
    optr = out.ptr + out_pad * np.dtype(dt).itemsize
    fn = {"ref": "bs_price_f32", "fast": "bs_price_fast_f32",
          "i32": "bs_price_i32type_f32", "f64": "bs_price_f64"}[mode]
    call(fn, *dbook.ptrs(), optr, dbook.n, stream)
    if sync:
        call("bs_stream_sync", stream)
    return out


def greeks(dbook: DeviceBook, outs=None, stream=None, sync: bool = True):
    if outs is None:
        outs = [alloc_out(dbook.n) for _ in range(5)]
    call("bs_price_greeks_f32", *dbook.ptrs(), *[o.ptr for o in outs], dbook.n, stream)
    if sync:
        call("bs_stream_sync", stream)
    return outs
