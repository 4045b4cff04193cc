"""The C ABI library without a GPU: it loads, exports every function the
header declares, and rejects bad arguments before any device work."""
import ctypes
import subprocess

import numpy as np

from p3arsec_amd import capi


def test_exports_every_declared_function():
    L = capi.load()
    declared = capi.header_functions()
    assert len(declared) >= 40
    missing = [f for f in declared if not hasattr(L, f)]
    assert not missing, missing
    # and the binding knows the signature of each of them
    assert sorted(capi._SIGS) == declared


def test_exported_symbols_are_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for f in capi.header_functions():
        assert f in names, f          # unmangled extern "C" names


def test_status_strings():
    L = capi.load()
    for code in capi.STATUS:
        assert L.bs_status_string(code)
    assert L.bs_status_string(12345) == b"unknown status"
    assert L.bs_version() >= 10000


def test_config_roundtrip_and_validation():
    L = capi.load()
    old = capi.get_config(capi.CFG_VECTOR_WIDTH)
    capi.set_config(capi.CFG_VECTOR_WIDTH, 2)
    assert capi.get_config(capi.CFG_VECTOR_WIDTH) == 2
    capi.set_config(capi.CFG_VECTOR_WIDTH, old)
    assert L.bs_config_set(capi.CFG_VECTOR_WIDTH, 3) == -1
    assert capi.get_config(capi.CFG_BLOCK_THREADS) == 0        # each kernel's default
    capi.set_config(capi.CFG_BLOCK_THREADS, 64)
    assert capi.get_config(capi.CFG_BLOCK_THREADS) == 64
    assert L.bs_config_set(capi.CFG_BLOCK_THREADS, 100) == -1  # 0, 64, 128 or 256
    assert L.bs_config_set(capi.CFG_BLOCK_THREADS, 512) == -1
    capi.set_config(capi.CFG_BLOCK_THREADS, 0)
    assert capi.get_config(capi.CFG_LDS_PER_WAVE) == -1        # each kernel's default
    capi.set_config(capi.CFG_LDS_PER_WAVE, 4096)
    assert capi.get_config(capi.CFG_LDS_PER_WAVE) == 4096
    assert L.bs_config_set(capi.CFG_LDS_PER_WAVE, 65537) == -1
    assert L.bs_config_set(capi.CFG_LDS_PER_WAVE, -2) == -1
    capi.set_config(capi.CFG_LDS_PER_WAVE, -1)
    assert L.bs_config_set(999, 1) == -1


def test_argument_errors_before_device_work():
    L = capi.load()
    nul = [None] * 7
    assert L.bs_price_f32(*nul, -1, None) == -1           # n < 0
    assert L.bs_price_f32(*nul, 0, None) == 0             # empty map: nothing to do
    assert L.bs_price_f32(*nul, 16, None) == -1           # NULL arrays
    assert b"NULL" in L.bs_last_error()
    assert L.bs_price_greeks_f32(*([None] * 11), 0, None) == 0
    # the allocation-free portfolio form holds books of at most 2^25 options
    assert L.bs_portfolio_f32_noalloc(*nul, (1 << 25) + 1, 0, None, None) == -1
    assert b"2^25" in L.bs_last_error()
    assert L.bs_price_host_f32(*nul, 8, 1, 0) == -1
    host = np.zeros(16, np.float32)
    p = host.ctypes.data
    # plain host memory is never handed to a kernel
    assert L.bs_price_f32(p, p, p, p, p, p, p, 16, None) == -1
    assert L.bs_repack_aos_f32(p + 1, 4, p, p, p, p, p, p, None) == -1   # misaligned AoS
    six = [None] * 6
    assert L.bs_price_passes_f32(*six, 1, None, 16, 0, None) == -1      # passes < 1
    assert L.bs_price_passes_f32(*six, 3, None, 16, 2, None) == -1      # type width 1 or 4
    assert L.bs_price_passes_f32(*six, 1, None, 0, 100, None) == 0      # empty book
    assert L.bs_price_passes_f32(*six, 1, None, 16, 100, None) == -1    # NULL arrays
    assert L.bs_price_passes_f32(p, p, p, p, p, p, 1, p, 16, 2, None) == -1   # host memory
    assert L.bs_event_create(None) == -1 and L.bs_event_create_timing(None) == -1


def test_scalar_and_thread_compat_arguments():
    L = capi.load()
    assert L.bs_price_scalar_f32(1, 1, 0.1, 0.2, 1, 0, 0, None) == -1
    assert L.bs_thread_compat(None) == -1
    assert L.bs_bind_globals(-1, 1, *([None] * 7)) == -1
    assert L.bs_bind_globals(10, 0, *([None] * 7)) == -1
    tid = ctypes.c_int(5)
    assert L.bs_bind_globals(10, 2, *([None] * 7)) == 0
    assert L.bs_thread_compat(ctypes.byref(tid)) == -1       # tid >= nthreads


def test_swaptions_argument_errors_before_device_work():
    from p3arsec_amd import swaptions as W
    L = capi.load()
    sws, _ = W.parsec_portfolio(2)

    def rc(sws_, trials=1000, seed=5, block=16, n=None):
        recs, keep = W._records(sws_)
        p = np.empty(max(len(sws_), 1))
        e = np.empty(max(len(sws_), 1))
        r = L.bs_swaptions_host_f64(recs, len(sws_) if n is None else n, trials, seed, block,
                                    p.ctypes.data, e.ctypes.data, None)
        del keep
        return r

    def mod(**kw):
        s = dict(sws[0])
        s.update(kw)
        return [s]

    assert L.bs_swaptions_host_f64(None, 0, 10, 0, 16, None, None, None) == 0   # nothing to do
    assert L.bs_swaptions_host_f64(None, 2, 10, 0, 16, None, None, None) == -1
    assert rc(sws, n=-1) == -1
    assert rc(sws, trials=0) == -1
    assert rc(sws, block=0) == -1
    assert rc(sws, seed=-1) == -1
    assert rc(sws, seed=2 ** 42) == -1
    assert rc(sws, trials=2 ** 38) == -1               # counters past 2^42
    assert b"2^42" in L.bs_last_error()
    # one launch holds at most 2^24 blocks of 1024 paths: 2^34 + 1 trials of one
    # swaption need one block more (refused before any device work)
    assert rc(sws[:1], trials=2 ** 34 + 1) == -1
    assert b"2^24 blocks" in L.bs_last_error()
    assert rc(sws, trials=2 ** 33 + 1) == -1            # two swaptions: 2 x (2^23 + 1) blocks
    assert b"2^24 blocks" in L.bs_last_error()
    assert rc(mod(maturity=30.0)) == -1                 # start past the horizon
    assert b"horizon" in L.bs_last_error()
    assert rc(mod(interval=0.01)) == -1                 # payments under one step
    assert rc(mod(tenor=40.0)) == -1                    # swap past the curve
    assert rc(mod(years=float("nan"))) == -1
    assert rc(mod(years=0.0)) == -1
    one = mod(**{"yield": np.array([0.1])}, factors=np.zeros((3, 0)))
    assert rc(one) == -1                                # n_steps < 2
    assert rc(mod(factors=np.zeros((9, 10)))) == -1     # 9 factors
    assert rc([sws[0], dict(sws[1], factors=W.PARSEC_FACTORS[:2])]) == -1   # mixed n_factors
    assert b"same n_factors" in L.bs_last_error()
    long_ = mod(**{"yield": 0.05 + 0.001 * np.arange(40)}, factors=np.full((8, 39), 0.01),
                years=40.0, maturity=9.0, tenor=2.0)
    assert rc(long_) == -1                              # start 9 x 8 factors > 64 draws
