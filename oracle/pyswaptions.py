"""ctypes front of the swaptions oracle in liboracle.so (swaptions_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never uses
it.  A swaption is a dict with the fields of the reference's `parm` record
(parsec-ff/pkgs/apps/swaptions/src/HJM_Securities.cpp:275-296): strike,
compounding, maturity, tenor, interval (dPaymentInterval), years, yield
(n_steps values) and factors (n_factors x (n_steps - 1)).
"""
from __future__ import annotations

import ctypes

import numpy as np

import pyoracle

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_done = False


def lib():
    global _done
    L = pyoracle.lib()
    if not _done:
        L.swo_ranunif.argtypes = [_i64p]
        L.swo_ranunif.restype = ctypes.c_double
        L.swo_ranunif_direct.argtypes = [_i64]
        L.swo_ranunif_direct.restype = ctypes.c_double
        L.swo_minstd_iterate.argtypes = [_i64, _i64]
        L.swo_minstd_iterate.restype = _i64
        L.swo_cum_normal_inv.argtypes = [ctypes.c_double]
        L.swo_cum_normal_inv.restype = ctypes.c_double
        L.swo_yield_to_forward.argtypes = [_f64p, ctypes.c_int, _f64p]
        L.swo_drifts.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 _f64p]
        sw_args = ([_f64p] + [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int,
                   ctypes.c_double, _f64p, _f64p, _i64, _i64, ctypes.c_int])
        for f in (L.swo_swaption_blocking, L.swo_swaption_chain):
            f.argtypes = sw_args
            f.restype = ctypes.c_int
        L.swo_chain_payoffs.argtypes = ([ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, _f64p, _f64p, ctypes.c_void_p, _i64,
                                        ctypes.c_void_p])
        L.swo_chain_payoffs.restype = ctypes.c_int
        L.swo_portfolio.argtypes = [ctypes.c_int, _i64p, _i64p, _f64p, _f64p]
        L.swo_price_batch.argtypes = ([ctypes.c_int] + [_f64p] * 6 + [ctypes.c_int, ctypes.c_int,
                                      _f64p, _f64p, _i64, _i64, ctypes.c_int, ctypes.c_int,
                                      _f64p, _f64p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int)])
        L.swo_price_batch.restype = ctypes.c_double
        _done = True
    return L


def ranunif(counter: int) -> float:
    c = ctypes.c_int64(counter)
    return lib().swo_ranunif(ctypes.byref(c))


def ranunif_direct(counter: int) -> float:
    return lib().swo_ranunif_direct(counter)


def minstd_iterate(x: int, steps: int) -> int:
    return lib().swo_minstd_iterate(x, steps)


def cum_normal_inv(u) -> np.ndarray:
    f = lib().swo_cum_normal_inv
    return np.array([f(float(v)) for v in np.ravel(u)])


def yield_to_forward(yld) -> np.ndarray:
    y = np.ascontiguousarray(yld, np.float64)
    out = np.empty_like(y)
    lib().swo_yield_to_forward(out, len(y), y)
    return out


def drifts(years: float, factors) -> tuple[np.ndarray, np.ndarray]:
    """(total drift (n-1), per-factor drifts (F x (n-1)))."""
    fa = np.ascontiguousarray(factors, np.float64)
    F, w = fa.shape
    total = np.empty(w)
    per = np.empty((F, w))
    lib().swo_drifts(total, per, w + 1, F, years, fa)
    return total, per


def _args(sw):
    y = np.ascontiguousarray(sw["yield"], np.float64)
    fa = np.ascontiguousarray(sw["factors"], np.float64)
    return (float(sw["strike"]), float(sw["compounding"]), float(sw["maturity"]),
            float(sw["tenor"]), float(sw["interval"]), len(y), fa.shape[0],
            float(sw["years"]), y, fa)


def price_blocking(sw, seed: int, trials: int, block: int = 16):
    """HJM_Swaption_Blocking (HJM_Swaption_Blocking.cpp:20-222): (price, stderr)."""
    out = np.empty(2)
    ok = lib().swo_swaption_blocking(out, *_args(sw), seed, trials, block)
    if not ok:
        raise ValueError("swaption sizes the reference cannot price")
    return float(out[0]), float(out[1])


def price_chain(sw, seed: int, trials: int, block: int = 16):
    out = np.empty(2)
    ok = lib().swo_swaption_chain(out, *_args(sw), seed, trials, block)
    if not ok:
        raise ValueError("swaption sizes the reference cannot price")
    return float(out[0]), float(out[1])


def start_step(sw) -> int:
    """The swaption's start step m (draws per path: m x n_factors)."""
    m = lib().swo_chain_payoffs(*_args(sw), None, 0, None)
    if m < 0:
        raise ValueError("swaption sizes the reference cannot price")
    return m


def chain_payoffs(sw, z) -> np.ndarray:
    """Discounted payoff of each path from caller-supplied normal draws
    z[paths, m * n_factors] (the model apart from the generator)."""
    z = np.ascontiguousarray(z, np.float64)
    out = np.empty(z.shape[0])
    m = lib().swo_chain_payoffs(*_args(sw), z.ctypes.data, z.shape[0], out.ctypes.data)
    if m < 0:
        raise ValueError("swaption sizes the reference cannot price")
    return out


def portfolio(n: int, seed: int = 1979):
    """HJM_Securities.cpp:198,275-296: (base seed, years[n], strike[n], seed after)."""
    s = ctypes.c_int64(seed)
    base = ctypes.c_int64(0)
    years = np.empty(n)
    strike = np.empty(n)
    lib().swo_portfolio(n, ctypes.byref(s), ctypes.byref(base), years, strike)
    return int(base.value), years, strike, int(s.value)


def price_batch(sws, seed: int, trials: int, block: int = 16, nthreads: int = 0,
                chain: bool = False):
    """The FastFlow ParallelFor over swaptions (HJM_Securities.cpp:311-323),
    swaption i with seed + i.  All swaptions share n_steps and n_factors.
    chain=True prices each swaption by the per-chain form (only the time steps
    its payoff reads; the same bits as the blocked form).
    Returns (price[ns], stderr[ns], seconds, threads)."""
    ns = len(sws)
    n = len(sws[0]["yield"])
    F = np.asarray(sws[0]["factors"]).shape[0]
    col = {k: np.ascontiguousarray([float(s[k]) for s in sws])
           for k in ("strike", "compounding", "maturity", "tenor", "interval", "years")}
    y = np.ascontiguousarray(np.stack([np.asarray(s["yield"], np.float64) for s in sws]))
    fa = np.ascontiguousarray(np.stack([np.asarray(s["factors"], np.float64) for s in sws]))
    assert y.shape == (ns, n) and fa.shape == (ns, F, n - 1)
    price = np.empty(ns)
    se = np.empty(ns)
    used = ctypes.c_int(0)
    sec = lib().swo_price_batch(ns, col["strike"], col["compounding"], col["maturity"],
                                col["tenor"], col["interval"], col["years"], n, F, y,
                                fa.reshape(-1), seed, trials, block, int(chain), price, se,
                                nthreads,
                                ctypes.byref(used))
    if sec < 0:
        raise ValueError("swaption batch failed")
    return price, se, sec, used.value
