"""Swaptions (HJM Monte-Carlo) front over the C ABI.

A swaption is a dict with the reference's `parm` fields
(parsec-ff/pkgs/apps/swaptions/src/HJM_Securities.cpp:275-296): strike,
compounding, maturity, tenor, interval (dPaymentInterval), years, yield
(n_steps values) and factors (n_factors rows of n_steps - 1).

parsec_portfolio() is the reference driver's input generation
(HJM_Securities.cpp:198,230-296): the three-factor volatility table, the
yield curve 0.10, 0.105, ..., the horizons 5-20 years in quarter steps and
strikes 0.1-5.0 drawn from its generator, and the base seed of the swaptions.
price() runs every path of every swaption on the GPU (bs_swaptions_host_f64);
there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi

N_STEPS = 11          # iN (HJM_Securities.cpp:56)
N_FACTORS = 3         # iFactors (HJM_Securities.cpp:58)
BLOCK_SIZE = 16       # BLOCK_SIZE of PARSEC's HJM_type.h
DEFAULT_TRIALS = 102400   # DEFAULT_NUM_TRIALS of PARSEC's HJM_type.h
DEFAULT_SEED = 1979   # HJM_Securities.cpp:61

# HJM_Securities.cpp:231-264: the volatility of each factor per maturity step.
PARSEC_FACTORS = np.array([
    [.01] * 10,
    [.009048, .008187, .007408, .006703, .006065, .005488, .004966, .004493, .004066, .003679],
    [.001000, .000750, .000500, .000250, .000000, -.000250, -.000500, -.000750, -.001000,
     -.001250],
])

_M31 = 2147483647


def ran_unif(seed: list) -> float:
    """One draw of the reference's generator from counter seed[0] (advanced by
    one): MINSTD(1513517 c mod (2^31-1)) * 4.656612875e-10, for the driver's
    parameter draws on the host.  The reference's C remainders and sign
    correction end on the residue in [0, 2^31-1), which Python's % gives
    directly, negative counters included."""
    c = seed[0]
    seed[0] = c + 1
    x = (16807 * ((c * 1513517) % _M31)) % _M31
    return x * 4.656612875e-10


def parsec_portfolio(n: int, seed: int = DEFAULT_SEED):
    """(swaptions, base seed): the driver's n swaptions; swaption i is priced
    with base seed + i (HJM_Securities.cpp:198,275-296)."""
    s = [seed]
    base = int(_M31 * ran_unif(s))
    # the driver adds .005 step by step: keep its sums exactly
    yld = np.empty(N_STEPS)
    yld[0] = .1
    for j in range(1, N_STEPS):
        yld[j] = yld[j - 1] + .005
    out = []
    for _ in range(n):
        years = 5.0 + int(60 * ran_unif(s)) * 0.25
        strike = 0.1 + int(49 * ran_unif(s)) * 0.1
        out.append({"strike": strike, "compounding": 0.0, "maturity": 1.0, "tenor": 2.0,
                    "interval": 1.0, "years": years, "yield": yld.copy(),
                    "factors": PARSEC_FACTORS.copy()})
    return out, base


def _records(sws):
    keep = []
    recs = (capi.Swaption * len(sws))()
    for r, sw in zip(recs, sws):
        y = np.ascontiguousarray(sw["yield"], np.float64)
        fa = np.ascontiguousarray(sw["factors"], np.float64)
        if fa.ndim != 2 or fa.shape[1] != len(y) - 1:
            raise ValueError("factors must be n_factors x (n_steps - 1)")
        keep += [y, fa]
        r.strike, r.compounding = float(sw["strike"]), float(sw["compounding"])
        r.maturity, r.tenor = float(sw["maturity"]), float(sw["tenor"])
        r.payment_interval, r.years = float(sw["interval"]), float(sw["years"])
        r.n_steps, r.n_factors = len(y), fa.shape[0]
        r.yield_ = y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        r.factors = fa.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return recs, keep


class Prepared:
    """The C records of a list of swaptions, built once (the reference builds
    its parm array before its timed region, HJM_Securities.cpp:266-302)."""

    def __init__(self, sws):
        self.n = len(sws)
        self.recs, self._keep = _records(sws)


def price(sws, trials: int = DEFAULT_TRIALS, seed: int = 0, block_size: int = BLOCK_SIZE,
          stream=None):
    """(price[n], stderr[n]) of swaptions sws (a list of dicts or a Prepared),
    swaption i with seed + i."""
    prep = sws if isinstance(sws, Prepared) else Prepared(sws)
    p = np.empty(prep.n)
    e = np.empty(prep.n)
    capi.call("bs_swaptions_host_f64", prep.recs, prep.n, trials, seed, block_size,
              p.ctypes.data, e.ctypes.data, stream)
    return p, e
